// FP32 selection for CompositeCutter (compositecutter.cpp:58-194): two
// sub-cutters (CylCutter, BallCutter or BullCutter), each with its radius
// window and z offset.  Every triangle carries one FP32 selection record per
// sub-cutter (drop_fast.hpp's layout), built in the frame lowered by the
// sub-cutter's z offset, so the record's heights are already the composite's
// (sub-cutter height - zoffset).  The contacts of the reference's composite
// drop of one triangle (dropCutter, millingcutter.cpp:386-409):
//   - facet, per sub-cutter n in order: the sub-cutter's facet contact when
//     its CC point lies in window n (the XY distance of a facet contact from
//     the CL point is fixed per triangle and sub-cutter: the record holds
//     its window test);
//   - vertices (when no facet lifted): the composite profile, the
//     sub-cutter whose window holds the vertex's distance q (the first one,
//     radius_to_index);
//   - edges (when no facet lifted), per sub-cutter in order: the highest of
//     its three edge contacts, kept only when that one's CC point lies in
//     its window (the edge loop lifts a copy of the CL point, and only the
//     final CC point is tested).
// A contact whose window test, or whose place among its sub-cutter's edges,
// is within the FP32 error of changing is ambiguous, and so is a facet lift
// decided within error (the facet-first rule then depends on the running
// height): such points escalate to the exact FP64 fold, as for cones.  The
// evaluation runs in rank order (no floor, no reordering, no near-tie sets).
#pragma once

#include "drop_fast.hpp"

namespace ocl_amd {

enum : int {
    R_CWIN = 71,   // per sub-cutter record: its facet contact's window test (1 in, 0 out, 2 within error)
    R_CBZ1 = 78,   // record 0: max z (composite frame) + the height error bound
    R_CMAXZ = 79,  // record 0: max z (composite frame)
    R_CTOP = 78,   // record 1: bit j set when vertex j is at the triangle's max z
};

// The composite selection takes two sub-cutters of kind Cyl, Ball or Bull,
// each at least as wide as its radius window.  (A sub-cutter narrower than
// its window gets the reference's beyond-radius height, -1 or NaN, on a band
// of vertices, which makes the composite's contacts order-dependent far more
// often: such composites keep the exact fold.)
OCL_HD bool comp_fast(const CutterDev& cd) {
    if (cd.c.kind != K_COMPOSITE || cd.nsub != 2) return false;
    for (int n = 0; n < 2; ++n) {
        const int k = cd.sub[n].kind;
        if (k != K_CYL && k != K_BALL && k != K_BULL) return false;
        if (cd.sub[n].radius < cd.rlim[n] - 1e-9) return false;
    }
    return cd.rlim[0] < cd.rlim[1];
}

// Radius window of sub-cutter n with the reference's 1e-6 tolerance
// (compositecutter.cpp:124-147).
OCL_HD double comp_lo(const CutterDev& cd, int n) { return n == 0 ? -1E-6 : cd.rlim[n - 1] - 1E-6; }
OCL_HD double comp_hi(const CutterDev& cd, int n) { return cd.rlim[n] + 1e-6; }

// Selection constants of a composite: the composite-level ones (radius,
// error bound, culling) in the returned SelParams and each sub-cutter's in
// sub[0..1]; the culling fields carry the windows and profiles.
template <class T>
OCL_HD SelParams<T> comp_sel_params(const CutterDev& cd, const SelErr& er, SelParams<T>* sub) {
    SelParams<T> P{};
    double E = 0.0, epos = 0.0;
    for (int n = 0; n < 2; ++n) {
        CutterDev s{};
        s.c = cd.sub[n];
        sub[n] = sel_params<T>(s, er);
        E = fmax(E, (double)sub[n].E);
        epos = fmax(epos, (double)sub[n].EPOS);
        P.ck[n] = cd.sub[n].kind;
        P.cR[n] = (T)cd.sub[n].radius;
        P.cr1[n] = (T)cd.sub[n].r1;
        P.cr2[n] = (T)cd.sub[n].r2;
        P.clo[n] = (T)comp_lo(cd, n);
        P.chi[n] = (T)comp_hi(cd, n);
        // a CylCutter's edge contacts lie on its rim, at XY distance R from
        // the CL point up to FP64 rounding: their window test is fixed
        const double R = cd.sub[n].radius, lo = comp_lo(cd, n), hi = comp_hi(cd, n);
        const double eq = 1e-12 * (1.0 + R);
        P.cwin[n] = (fabs(R - lo) <= eq || fabs(R - hi) <= eq) ? 2 : ((R >= lo && R <= hi) ? 1 : 0);
        P.cflat[n] = (cd.sub[n].kind == K_CYL || cd.sub[n].kind == K_BULL) && cd.zoff[n] == 0.0;
    }
    const double R = cd.c.radius;
    P.kind = K_COMPOSITE;
    P.R = (T)R;
    P.R2 = (T)(R * R);
    P.DM = (T)er.dm();
    P.E = (T)E;
    P.EPOS = (T)epos;
    P.SQDM = (T)sqrt(2.0 * er.dm());
    P.HPAD = (T)(16.0 * er.u * R);
    P.CUT = (T)INFINITY;
    P.cull = 1;
    return P;
}

// Lower bound of the composite profile at XY distance sqrt(d2) (rounded
// down): a contact of sub-cutter n lies at a distance in window n, where
// its profile height is at least h_n(max(d, lo_n)); the record frames hold
// the z offsets.  +inf when no window reaches d.
template <class T>
OCL_HD T comp_prof_lower(T d2, const SelParams<T>& P) {
    typedef FM<T> F;
    const T d = F::sqrt_(d2);
    T h = F::inf();
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const T r = F::max_(d, P.clo[n]);
        T hn = (T)0;
        if (P.ck[n] == K_BALL) {
            hn = P.cR[n] - F::sqrt_(F::max_(P.cR[n] * P.cR[n] - r * r, (T)0));
        } else if (P.ck[n] == K_BULL) {
            const T w = F::max_(r - P.cr1[n], (T)0);
            hn = P.cr2[n] - F::sqrt_(F::max_(P.cr2[n] * P.cr2[n] - w * w, (T)0));
        }
        h = F::min_(h, d <= P.chi[n] ? F::max_(hn - P.HPAD, (T)0) : F::inf());
        // beyond a CylCutter's or BullCutter's radius, inside its window's
        // tolerance, the reference's height is -1 (comp_vertex_sentinel)
        const bool sent = (P.ck[n] != K_BALL) & (d <= P.chi[n] + P.EPOS);
        h = F::min_(h, sent ? (T)-1 - P.HPAD : F::inf());
    }
    return h;
}

// Upper bound of every contact's upper value of the triangle (records rc:
// sub-cutter 0's, then sub-cutter 1's) for a lane at (x, y) whose distance
// to the triangle's box is sqrt(d2): per sub-cutter reaching d, the lower of
// its profile cap and its plane cap.
template <class T>
OCL_HD T comp_tri_cap(const T* rc, T x, T y, T d2, const SelParams<T>& P) {
    typedef FM<T> F;
    const T d = F::sqrt_(d2);
    T cap = -F::inf();
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const T* r = rc + rec_len<T>() * n;
        const T rr = F::max_(d, P.clo[n]);
        T hn = (T)0;
        if (P.ck[n] == K_BALL) {
            hn = P.cR[n] - F::sqrt_(F::max_(P.cR[n] * P.cR[n] - rr * rr, (T)0));
        } else if (P.ck[n] == K_BULL) {
            const T w = F::max_(rr - P.cr1[n], (T)0);
            hn = P.cr2[n] - F::sqrt_(F::max_(P.cr2[n] * P.cr2[n] - w * w, (T)0));
        }
        const T zp = r[R_ZCAP] - F::max_(hn - P.HPAD, (T)0);
        const T pl = F::fma_(r[R_GY], y, F::fma_(r[R_GX], x, r[R_F0])) + r[R_PMAR];
        cap = F::max_(cap, d <= P.chi[n] ? F::min_(zp, pl) : -F::inf());
        // vertex contacts at the reference's -1 height beyond the
        // sub-cutter's radius (comp_vertex_sentinel): above both caps
        if (P.ck[n] != K_BALL && d <= P.chi[n] + P.EPOS) {
            const T base = P.ck[n] == K_BULL ? P.cr2[n] : (T)0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const T dx = rc[R_P0X + 3 * j] - x, dy = rc[R_P0Y + 3 * j] - y;
                const T q = F::sqrt_(F::fma_(dx, dx, dy * dy));
                const bool band = (q >= P.cR[n] - P.EPOS) & (q <= P.chi[n] + P.EPOS);
                cap = F::max_(cap, band ? r[R_P0Z + 3 * j] + base + (T)1 + r[R_EMAX] + P.E : -F::inf());
            }
        }
    }
    return cap;
}

// Window test of a CC point at XY distance q from the CL point for
// sub-cutter n: 1 in, 0 out, 2 within eq of a window end.
template <class T>
OCL_HD int comp_window(T q, int n, const SelParams<T>& P, T eq) {
    const bool in = (q >= P.clo[n]) & (q <= P.chi[n]);
    const bool near = (fabs(q - P.clo[n]) <= eq) | (fabs(q - P.chi[n]) <= eq);
    return near ? 2 : (in ? 1 : 0);
}

// Builds the two sub-cutter records of a triangle (FP32 only): record n in
// the frame (ox, oy, oz + zoff[n]); the facet's window test in R_CWIN of
// each, the composite-frame max z in R_CBZ1 of record 0.
template <class T>
OCL_HD void build_record_comp(const TriX& t, const CutterDev& cd, double ox, double oy, double oz, const SelErr& er,
                              T* out, const uint32_t* fid) {
    const double Eabs = er.E();
    const V3 nrm = up_normal(t);
    const bool vertical = near0(nrm.z);
    const bool horiz = !vertical && near0(nrm.x) && near0(nrm.y);
    const V3 nn = vunit(nrm);
    const double nxy = sqrt(nn.x * nn.x + nn.y * nn.y);
    for (int n = 0; n < 2; ++n) {
        CutterDev s{};
        s.c = cd.sub[n];
        T* r = out + rec_len<T>() * n;
        build_record<T>(t, s, ox, oy, oz + cd.zoff[n], er, r, fid);
        // sub-cutter edge IDs: a contact of the other sub-cutter on the same
        // edge is a different contact
        if constexpr (sizeof(T) == 4) {
            if (n == 1)
                for (int j = 0; j < 3; ++j)
                    r[R_FID0 + 3 + j] = (T)id_bits(id_of((float)r[R_FID0 + 3 + j]) | 0x20000000u);
        }
        // XY distance of the facet contact from the CL point: |xyl xyn +
        // nl n|_xy (millingcutter.cpp:51-75), 0 on a horizontal facet
        const double q = horiz ? 0.0 : s.c.xyl + s.c.nl * nxy;
        const double lo = comp_lo(cd, n), hi = comp_hi(cd, n);
        const double eq = 1e-12 * (1.0 + cd.c.radius);
        const bool in = !(q < lo) && !(q > hi);
        const bool near = fabs(q - lo) <= eq || fabs(q - hi) <= eq;
        r[R_CWIN] = (T)(near ? 2.0 : (in ? 1.0 : 0.0));
    }
    static_assert(sizeof(T) == 4 || R_CBZ1 >= kFastRec, "composite records are FP32");
    if constexpr (sizeof(T) == 4) {
        out[R_CBZ1] = (T)(t.maxz - oz + Eabs);
        out[R_CMAXZ] = (T)(t.maxz - oz);
        int top = 0;
        for (int j = 0; j < 3; ++j) top |= (tri_vtx(t, j).z == t.maxz) ? (1 << j) : 0;
        out[rec_len<T>() + R_CTOP] = (T)top;
    }
}

// The edge contact one sub-cutter contributes: the highest of its three
// (strict '>', the first of equal heights), when its CC point lies in the
// sub-cutter's window; its other edge contacts never count.  amb: which
// contact is highest, whether it is valid, or its window test, is within
// error.
template <class T>
struct CompEdges {
    T z, e, a;    // height, error bound, slack of the contact (z = -inf: none)
    T u;          // the largest upper value of a contact that may be the contribution
    bool amb;
    int j;        // its edge
    uint32_t id;  // its feature ID
};

template <class T, int K>
OCL_HD CompEdges<T> comp_edges(const T* r, int n, T x, T y, const SelParams<T>& Ps, const SelParams<T>& P,
                               bool ve) {
    typedef FM<T> F;
    const T NEG = -F::inf();
    const int flags = (int)r[R_FLAGS];
    EdgeC<T> ec[3];
    int w[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        ec[j] = edge_contact<T, K>(r, j, flags, x, y, Ps, ve);
        if (K == K_CYL) {
            w[j] = P.cwin[n];
        } else {
            const T q = F::sqrt_(F::fma_(ec[j].sd, ec[j].sd, ec[j].ccu * ec[j].ccu));
            w[j] = comp_window<T>(q, n, P, Ps.EPOS + Ps.SQDM);
        }
    }
    T zmax = NEG, emax = (T)0, amax = (T)0;
    int jm = -1, wm = 0;
    bool am = false;
    uint32_t idm = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const bool take = ec[j].val & (ec[j].z > zmax);
        zmax = take ? ec[j].z : zmax;
        emax = take ? ec[j].eze : emax;
        amax = take ? ec[j].eae : amax;
        am = take ? ec[j].amb : am;
        wm = take ? w[j] : wm;
        jm = take ? j : jm;
        idm = take ? id_of((float)r[R_FID0 + 3 + j]) : idm;
    }
    bool amb = (jm >= 0) & (am | (wm == 2));
    T umax = jm >= 0 ? zmax + F::max_(emax, amax) : NEG;
    // another contact that may be the highest (valid within error, or
    // ambiguous) and counts when it is: its upper value joins
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T u = ec[j].val ? ec[j].z + ec[j].eze : (ec[j].amb ? ec[j].z + ec[j].eae : NEG);
        const bool rival = (j != jm) & (ec[j].val | ec[j].amb) & (u >= zmax - emax) & (w[j] != 0) &
                           ((w[j] != wm) | (jm < 0));
        amb = amb | rival;
        umax = F::max_(umax, rival ? u : NEG);
    }
    const bool keep = (jm >= 0) & (wm != 0);
    CompEdges<T> o;
    o.z = keep ? zmax : NEG;
    o.e = emax;
    o.a = amax;
    o.amb = amb;
    o.u = umax;
    o.j = jm;
    o.id = idm;
    return o;
}

// No edge contact (what comp_edges gives a lane with ve = false).
template <class T>
OCL_HD CompEdges<T> comp_edges_none() {
    CompEdges<T> o;
    o.z = -FM<T>::inf();
    o.e = (T)0;
    o.a = (T)0;
    o.u = -FM<T>::inf();
    o.amb = false;
    o.j = -1;
    o.id = 0;
    return o;
}

// comp_edges for sub-cutter n's kind; skipped (no contact) when no lane of
// the wavefront needs it.
template <class T>
OCL_HD CompEdges<T> comp_edges_kind(const T* r, int n, T x, T y, const SelParams<T>& Ps, const SelParams<T>& P,
                                    bool ve) {
    if (!sel_any(ve)) return comp_edges_none<T>();
    if (P.ck[n] == K_CYL) return comp_edges<T, K_CYL>(r, n, x, y, Ps, P, ve);
    if (P.ck[n] == K_BALL) return comp_edges<T, K_BALL>(r, n, x, y, Ps, P, ve);
    return comp_edges<T, K_BULL>(r, n, x, y, Ps, P, ve);
}

// Vertex contact height of sub-cutter kind k (runtime) from record r.
template <class T>
OCL_HD T comp_vertex_height(const T* r, int j, int k, T x, T y, const SelParams<T>& Ps, T& m) {
    if (k == K_CYL) return vertex_height<T, K_CYL>(r, j, x, y, Ps, m);
    if (k == K_BALL) return vertex_height<T, K_BALL>(r, j, x, y, Ps, m);
    return vertex_height<T, K_BULL>(r, j, x, y, Ps, m);
}

// One candidate triangle for a composite cutter (FP32): as sel_triangle,
// with the contacts listed at the top of this file.  rc: the triangle's two
// records; P: composite constants, PS: the sub-cutters'.
template <class T>
OCL_HD void sel_triangle_comp(const T* rc, uint32_t k, T x, T y, const SelParams<T>& P, const SelParams<T>* PS,
                              Sel<T>& S, bool on, T* hi = nullptr) {
    typedef FM<T> F;
    const T NEG = -F::inf();
    const T* r0 = rc;
    const T* r1 = rc + rec_len<T>();
    const bool live = on & (S.zb < r0[R_CBZ1]);
    S.pairs += live ? 1u : 0u;
    // ---------------- facets, sub-cutter 0 then 1
    T zfac = NEG, efac = (T)0, th = NEG, uf[2];
    int nfac = -1;
    bool fambw = false, fvalany = false;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const T* r = n == 0 ? r0 : r1;
        const T u = F::fma_(r[R_U2], y, F::fma_(r[R_U1], x, r[R_U0]));
        const T v = F::fma_(r[R_V2], y, F::fma_(r[R_V1], x, r[R_V0]));
        const T mf = F::min_(F::min_(u, v), (T)1 - u - v);
        const T zf = F::fma_(r[R_GY], y, F::fma_(r[R_GX], x, r[R_F0]));
        const T ezf = r[R_EZF];
        const int w = (int)r[R_CWIN];
        const bool fval = live & (mf > (T)0) & (w != 0);
        const bool famb = live & (((fabs(mf) <= r[R_EBF]) & (w != 0)) | ((mf > -r[R_EBF]) & (w == 2)));
        uf[n] = fval ? zf + ezf : NEG;
        S.zt[2] = F::max_(S.zt[2], uf[n]);
        S.za = F::max_(S.za, famb ? zf + ezf : NEG);
        th = F::max_(th, (fval | famb) ? zf + ezf : NEG);
        fambw = fambw | famb;
        fvalany = fvalany | fval;
        const bool take = fval & (zf > zfac);
        zfac = take ? zf : zfac;
        efac = take ? ezf : efac;
        nfac = take ? n : nfac;
    }
    const bool flift = (nfac >= 0) & (zfac > S.zb);
    // the facet-first rule skips the vertices and edges only when a facet
    // lifts: decided within error, the result depends on the order
    S.oamb = S.oamb | (fvalany & (fabs(zfac - S.zb) <= efac + S.eb));
    const bool ve = live & !flift;
    // ---------------- vertices: the sub-cutter whose window holds q
    T zv = NEG, upv[3];
    int fv = FT_NONE;
    bool vamb = false;
    // a valid vertex contact at exactly the triangle's max z (its vertex is
    // the highest and its profile height is exactly 0) lifts the reference
    // to at least max z, so its edges are surely skipped
    bool top_exact = false;
    const int topm = (int)r1[R_CTOP];
    const T eqv = P.EPOS;  // error bound of a vertex distance q
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const T dx = r0[R_P0X + 3 * j] - x, dy = r0[R_P0Y + 3 * j] - y;
        const T q2 = F::fma_(dx, dx, dy * dy);
        const T q = F::sqrt_(q2);
        const T mc = P.R2 - q2;  // under the composite: q <= its radius
        // the sub-cutter whose window holds q (the first, radius_to_index)
        const int n = q <= P.chi[0] ? 0 : 1;
        const bool wamb = fabs(q - P.chi[0]) <= eqv;  // or the other one
        // Per sub-cutter: its profile height and, beyond its own radius but
        // inside its window's 1e-6 tolerance, the reference's height there:
        // -1 for CylCutter and BullCutter, NaN (no contact) for BallCutter
        // (prof_height; cylcutter.hpp, bullcutter.cpp, ballcutter.hpp
        // height()), which puts the contact at the vertex height + 1.
        T zc[2], zx[2], ms[2];
        bool bey[2], nr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const T* r = i == 0 ? r0 : r1;
            zc[i] = comp_vertex_height<T>(r, j, P.ck[i], x, y, PS[i], ms[i]);
            const T base = P.ck[i] == K_BALL ? P.cR[i] : (P.ck[i] == K_BULL ? P.cr2[i] : (T)0);
            zx[i] = r[R_P0Z + 3 * j] + base + (T)1;
            bey[i] = q > P.cR[i] + eqv;
            nr[i] = fabs(q - P.cR[i]) <= eqv;
        }
        const bool bn = n == 0 ? bey[0] : bey[1];
        const int kn = n == 0 ? P.ck[0] : P.ck[1];
        const T z = bn ? (n == 0 ? zx[0] : zx[1]) : (n == 0 ? zc[0] : zc[1]);
        const T mn = n == 0 ? ms[0] : ms[1];
        const T s2v = n == 0 ? PS[0].S2V : PS[1].S2V, eav = n == 0 ? PS[0].EAV : PS[1].EAV;
        const bool val = ve & (mc >= (T)0) & !(bn & (kn == K_BALL));
        // every height the reference may take here when a test is within error
        bool unsure = (mn <= s2v) & !bn;
        T zamb = val ? z : NEG;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bool may = (n == i) | wamb;
            const bool sure_bey = bey[i] & !nr[i];
            const T zi = sure_bey ? (P.ck[i] == K_BALL ? NEG : zx[i]) : zc[i];
            zamb = F::max_(zamb, may ? zi : NEG);
            zamb = F::max_(zamb, (may & nr[i] & (P.ck[i] != K_BALL)) ? zx[i] : NEG);
            unsure = unsure | (may & nr[i]);
        }
        unsure = unsure | wamb;
        const bool amb = ve & (((mc <= P.DM) & (mc >= -P.DM)) | ((mc >= (T)0) & unsure));
        const bool flat0 = (n == 0 ? P.cflat[0] : P.cflat[1]) != 0;
        const bool zero = flat0 & ((kn == K_CYL) | ((kn == K_BULL) & (q <= (n == 0 ? P.cR[0] - P.cr2[0] : P.cR[1] - P.cr2[1]) - eqv)));
        top_exact = top_exact | (val & !amb & !bn & zero & (((topm >> j) & 1) != 0));
        const T up = val ? z + P.E : NEG;
        upv[j] = up;
        S.zt[0] = F::max_(S.zt[0], up);
        const T za = zamb + eav;
        S.za = F::max_(S.za, amb ? za : NEG);
        th = F::max_(th, F::max_(up, amb ? za : NEG));
        const bool take = val & (z > zv);
        zv = take ? z : zv;
        fv = take ? (FT_VTX0 + j) : fv;
        vamb = take ? amb : vamb;
    }
    // ---------------- edges: per sub-cutter, its highest contact in window,
    // only while the height after the vertices is below the triangle's max
    // z (millingcutter.cpp:386-409: a composite's vertex contact can reach
    // above it when a z offset is negative); decided within error, the
    // result depends on the order
    const T zafter = F::max_(S.zb, zv);
    const T cmaxz = r0[R_CMAXZ];
    const bool eon = ve & (zafter < cmaxz) & !top_exact;
    const CompEdges<T> E0 = comp_edges_kind<T>(r0, 0, x, y, PS[0], P, eon);
    const CompEdges<T> E1 = comp_edges_kind<T>(r1, 1, x, y, PS[1], P, eon);
    {
        // When that test is within error, the edges the reference may or may
        // not run join as ambiguous competitors (a tie with the running
        // height is then decided exactly), and one clearly above the running
        // height makes the result order-dependent.
        const bool close = ve & !top_exact & (fabs(zafter - cmaxz) <= P.E + P.E);
        const CompEdges<T> A0 = comp_edges_kind<T>(r0, 0, x, y, PS[0], P, close & !eon);
        const CompEdges<T> A1 = comp_edges_kind<T>(r1, 1, x, y, PS[1], P, close & !eon);
        const T ua = F::max_(F::max_(A0.u, A1.u), F::max_(A0.z + A0.e, A1.z + A1.e));
        S.za = F::max_(S.za, ua);
        th = F::max_(th, ua);
        const T ze = F::max_(F::max_(A0.z, A1.z), F::max_(E0.z, E1.z));
        S.oamb = S.oamb | (close & (ze > zafter + P.E + P.E));
    }
    const T ue0 = E0.z + E0.e, ue1 = E1.z + E1.e;  // -inf when none
    S.zt[1] = F::max_(S.zt[1], F::max_(ue0, ue1));
    th = F::max_(th, F::max_(ue0, ue1));
    // an uncertain contribution (validity, window, or which contact is the
    // sub-cutter's highest) can only be a contact up to its upper value
    S.za = F::max_(S.za, E0.amb ? E0.u : NEG);
    S.za = F::max_(S.za, E1.amb ? E1.u : NEG);
    th = F::max_(th, F::max_(E0.amb ? E0.u : NEG, E1.amb ? E1.u : NEG));
    // vertices first, then sub-cutter 0's edges, then sub-cutter 1's;
    // strict '>' keeps the first of equal heights
    const bool e0w = E0.z > zv;
    const T zve = e0w ? E0.z : zv;
    const bool e1w = E1.z > zve;
    const T zw = e1w ? E1.z : zve;
    const T ew = e1w ? E1.e : (e0w ? E0.e : P.E);
    const int fw = e1w ? ((1 << 4) | (FT_EDGE0 + E1.j)) : (e0w ? (FT_EDGE0 + E0.j) : fv);
    const bool aw = e1w ? E1.amb : (e0w ? E0.amb : vamb);
    const bool wlift = (fw != FT_NONE) & (zw > S.zb);
    const bool any = flift | wlift;
    // runner-up (as sel_triangle): the other valid contacts of the triangle
    // when it lifts, else those that are not copies of the winner's feature
    if (sel_any(live & (th >= S.zb - P.CUT))) {
        uint32_t idw = 0;
        T uexc = NEG, unew = NEG;
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            const bool win = flift & (nfac == n);
            uexc = F::max_(uexc, win ? NEG : uf[n]);
            unew = F::max_(unew, uf[n]);
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const uint32_t iv = id_of((float)r0[R_FID0 + j]);
            const bool bv = !flift & !e0w & !e1w & (fv == FT_VTX0 + j);
            uexc = F::max_(uexc, (bv || flift) ? NEG : upv[j]);
            unew = F::max_(unew, iv != S.wid ? upv[j] : NEG);
            idw = bv ? iv : idw;
        }
        // the sub-cutters' edge contacts
        const bool b0 = !flift & e0w & !e1w, b1 = !flift & e1w;
        uexc = F::max_(uexc, (b0 || flift) ? NEG : ue0);
        uexc = F::max_(uexc, (b1 || flift) ? NEG : ue1);
        unew = F::max_(unew, E0.id != S.wid ? ue0 : NEG);
        unew = F::max_(unew, E1.id != S.wid ? ue1 : NEG);
        idw = b0 ? E0.id : (b1 ? E1.id : idw);
        const uint32_t wnew = flift ? facet_id(k, nfac == 1) : idw;
        const T uprev = (S.bt != 0xffffffffu && S.wid != wnew) ? S.zb + S.eb : NEG;
        S.zr = F::max_(S.zr, any ? F::max_(uprev, uexc) : unew);
        S.wid = any ? wnew : S.wid;
    }
    const T zn = flift ? zfac : zw;
    const T en = flift ? efac : ew;
    const int fn = flift ? ((nfac << 4) | FT_FACET) : fw;
    const bool an = flift ? false : aw;
    S.zb = any ? zn : S.zb;
    S.eb = any ? en : S.eb;
    S.bt = any ? k : S.bt;
    S.bf = any ? fn : S.bf;
    S.bamb = any ? an : S.bamb;
    if (hi) *hi = th;
}

}  // namespace ocl_amd
