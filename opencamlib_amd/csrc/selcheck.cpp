// Host build of the selection-pass arithmetic (drop_fast.hpp), for CPU tests
// of the product's host-compilable logic: record construction, the per-pair
// selection, the escalation decision and the exact FP64 refine.  It emulates
// one GPU lane per CL point (the lane's walk visits the level-0 boxes that
// overlap the point's own cutter box, in rank order).  Test support only:
// the product path always runs the HIP kernels.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/ocl_hip.h"
#include "cutter.hpp"
#include "drop_exact.hpp"
#include "drop_comp.hpp"
#include "drop_fast.hpp"
#include "feature_ids.hpp"
#include "kd_rank.hpp"
#include "push_exact.hpp"

using namespace ocl_amd;

namespace {

constexpr int kLeafH = 16;  // triangles per level-0 box (same as mesh.hpp kLeaf)

bool cutter_of(const ocl_cutter_desc* d, CutterDev& cd) {
    std::memset(&cd, 0, sizeof(cd));
    auto simple = [](const ocl_cutter_simple& s, CutterP& c) {
        switch (s.type) {
        case OCL_CUTTER_CYL: c = CutterP::cyl(s.diameter, s.length); return true;
        case OCL_CUTTER_BALL: c = CutterP::ball(s.diameter, s.length); return true;
        case OCL_CUTTER_BULL: c = CutterP::bull(s.diameter, s.corner_radius, s.length); return true;
        case OCL_CUTTER_CONE: c = CutterP::cone(s.diameter, s.angle, s.length); return true;
        default: return false;
        }
    };
    if (d->c.type == OCL_CUTTER_COMPOSITE) {  // as the library's cutter_from_desc
        if (d->nsub < 1 || d->nsub > kMaxSub) return false;
        cd.c.kind = K_COMPOSITE;
        cd.nsub = d->nsub;
        double r = 0;
        for (int i = 0; i < d->nsub; ++i) {
            if (!simple(d->sub[i], cd.sub[i])) return false;
            cd.rlim[i] = d->radius_limit[i];
            cd.hlim[i] = d->height_limit[i];
            cd.zoff[i] = d->zoffset[i];
            r = std::max(r, d->radius_limit[i]);
        }
        cd.c.radius = r;
        cd.c.diameter = 2 * r;
        cd.c.length = d->c.length;
        return true;
    }
    return simple(d->c, cd.c);
}

template <class T>
SelParams<T> params(const CutterDev& cd, const SelErr& er) {
    return sel_params<T>(cd, er);
}

template <class T, int K>
void run_points(const std::vector<TriX>& tx, const std::vector<BoxOrigin>& org, const std::vector<double>& bbx,
                const CutterDev& cd, const SelParams<T>& P, const SelParams<T>* PS, const double* cl, size_t n,
                const std::vector<T>& rec,
                const std::vector<T>& boxcap, const uint32_t* order, double* z_out, int* type_out, int* esc_out,
                int* feat_out, int* tri_out) {
    typedef FM<T> F;
    constexpr bool kComp = K == K_COMPOSITE;
    const size_t stride = (kComp ? 2 : 1) * (size_t)rec_len<T>();
    const double R = cd.c.radius;
    const size_t ntri = tx.size();
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)n; ++i) {
        const double x = cl[3 * i], y = cl[3 * i + 1], z0 = cl[3 * i + 2];
        Sel<T> S;
        sel_init(S, (T)0);
        T wzb = 0, wzt[4], wza = -F::inf(), wzr = -F::inf();
        for (int q = 0; q < 4; ++q) wzt[q] = -F::inf();
        bool framed = false;
        double ozw = 0;
        const size_t nbox = (ntri + kLeafH - 1) / kLeafH;
        // floor of the final height from the vertices of the boxes under the point
        T zc = -F::inf();
        if (P.cull && K != K_CONE && !kComp)
            for (size_t b = 0; b < nbox; ++b) {
                const double* bb = &bbx[4 * b];
                if (bb[1] < x || bb[0] > x || bb[3] < y || bb[2] > y) continue;
                const BoxOrigin& bo = org[b];
                if (!framed) {
                    framed = true;
                    ozw = bo.oz;
                    wzb = (T)(z0 - ozw);
                }
                const T dz = (T)(bo.oz - ozw);
                const T xl = (T)(x - bo.ox), yl = (T)(y - bo.oy);
                const size_t e = std::min(ntri, (b + 1) * kLeafH);
                for (size_t k = b * kLeafH; k < e; ++k)
                    zc = F::max_(zc, vertex_floor<T, K>(&rec[stride * k], xl, yl, P, true) + dz);
            }
        for (size_t b = 0; b < nbox; ++b) {
            const double* bb = &bbx[4 * b];
            if (bb[1] < x - R || bb[0] > x + R || bb[3] < y - R || bb[2] > y + R) continue;
            const BoxOrigin& bo = org[b];
            if (!framed) {
                framed = true;
                ozw = bo.oz;
                wzb = (T)(z0 - ozw);
            }
            const T dz = (T)(bo.oz - ozw);
            const T xl = (T)(x - bo.ox), yl = (T)(y - bo.oy);
            if (P.cull) {
                const T d2 = box_d2<T>(xl, yl, -(T)bo.hx, (T)bo.hx, -(T)bo.hy, (T)bo.hy);
                const T pl = kComp ? comp_prof_lower<T>(d2, P) : prof_lower<T, K>(d2, P);
                if (boxcap[b] + dz - pl < F::max_(wzb, zc) - P.CUT) continue;
            }
            const uint32_t bt_in = S.bt;
            S.zb = wzb - dz;
            S.zr = wzr - dz;
            for (int q = 0; q < 4; ++q) S.zt[q] = -F::inf();
            S.za = -F::inf();
            const size_t e = std::min(ntri, (b + 1) * kLeafH);
            for (size_t k = b * kLeafH; k < e; ++k) {
                const T* r = &rec[stride * k];
                if (r[R_BX1] < xl - P.R || r[R_BX0] > xl + P.R || r[R_BY1] < yl - P.R || r[R_BY0] > yl + P.R)
                    continue;
                if (P.cull) {
                    const T d2 = box_d2<T>(xl, yl, r[R_BX0], r[R_BX1], r[R_BY0], r[R_BY1]);
                    const T cap = kComp ? comp_tri_cap<T>(r, xl, yl, d2, P) : tri_cap<T, K>(r, xl, yl, d2, P);
                    if (cap < F::max_(S.zb, zc - dz) - P.CUT) continue;
                }
                if constexpr (kComp) sel_triangle_comp<T>(r, (uint32_t)k, xl, yl, P, PS, S, true);
                else sel_triangle<T, K>(r, (uint32_t)k, xl, yl, P, S, true);
            }
            if (S.bt != bt_in) wzb = S.zb + dz;
            wzr = S.zr + dz;
            for (int q = 0; q < 4; ++q) wzt[q] = F::max_(wzt[q], S.zt[q] + dz);
            wza = F::max_(wza, S.za + dz);
        }
        Sel<T> D = S;
        D.zb = wzb;
        for (int q = 0; q < 4; ++q) D.zt[q] = wzt[q];
        D.za = wza;
        D.zr = wzr;
        bool up = framed && sel_escalate(D, (T)(z0 - ozw), P);
        int why = 0;
        if (framed) {
            if (D.bamb) why |= 1;
            if (D.oamb) why |= 2;
            T other = -F::inf();
            const int sl = feat_slot(D.bf);
            for (int q = 0; q < 4; ++q)
                if (D.bt == 0xffffffffu || q != sl) other = F::max_(other, D.zt[q]);
            if (D.bt != 0xffffffffu) other = F::max_(other, (T)(z0 - ozw) + P.E);
            if (other >= D.zb - D.eb) why |= 4;
            if (D.za >= D.zb - D.eb) why |= 8;
        }
        double zr = z0;
        int ty = CC_NONE, ft = FT_NONE, tr = -1;
        if (D.bt != 0xffffffffu) {
            const TriX& t = tx[D.bt];
            DropLane X;
            X.x = x; X.y = y; X.z = -INFINITY; X.cc = v3(0, 0, 0); X.type = CC_NONE;
            X.tri = 0xffffffffu; X.feat = FT_NONE; X.cur_tri = D.bt; X.cur_feat = 0;
            int it = 0;
            if (kComp) exact_feature_comp(cd, X, t, D.bf, it);
            else exact_feature(cd.c, X, t, D.bf, it);
            zr = ozw + (double)D.zb;
            ty = X.type;
            ft = D.bf;
            tr = (int)order[D.bt];
            const double zgap = sizeof(T) == 4 ? (double)D.zr - (double)D.zb : -INFINITY;
            if (!up && (X.type == CC_NONE || feat_slot(X.feat) != feat_slot(D.bf) || !(X.z > z0) ||
                        fabs(X.z - zr) > 4.0 * ((double)D.eb + (double)P.E) || !(X.z > zr + zgap))) {
                up = true;
                why |= 16;
            }
            if (!up) zr = X.z;
        }
        z_out[i] = zr;
        type_out[i] = ty;
        esc_out[i] = up ? (1 | (why << 1)) : 0;
        feat_out[i] = ft;
        tri_out[i] = tr;
    }
}

}  // namespace

extern "C" int ocl_selcheck_drop(const double* tris, size_t ntri, unsigned bucket, const ocl_cutter_desc* desc,
                                 const double* cl, size_t n, int use_float, double escale, double* z_out,
                                 int* type_out, int* esc_out, int* feat_out, int* tri_out) {
    CutterDev cd;
    if (!cutter_of(desc, cd)) return OCL_E_UNSUPPORTED;
    std::vector<double> bb(6 * ntri);
    for (size_t i = 0; i < ntri; ++i)
        for (int d = 0; d < 3; ++d) {
            double lo = tris[9 * i + d], hi = lo;
            for (int k = 1; k < 3; ++k) {
                lo = std::fmin(lo, tris[9 * i + 3 * k + d]);
                hi = std::fmax(hi, tris[9 * i + 3 * k + d]);
            }
            bb[6 * i + 2 * d] = lo;
            bb[6 * i + 2 * d + 1] = hi;
        }
    static const int kXY[4] = {0, 1, 2, 3};
    const std::vector<uint32_t> order = kd_rank_order(bb.data(), ntri, kXY, bucket);
    std::vector<TriX> tx(ntri);
    for (size_t r = 0; r < ntri; ++r) {
        const double* q = tris + 9 * (size_t)order[r];
        TriX& t = tx[r];
        t.p[0] = v3(q[0], q[1], q[2]);
        t.p[1] = v3(q[3], q[4], q[5]);
        t.p[2] = v3(q[6], q[7], q[8]);
        t.n = vunit(vcross(vsub(t.p[0], t.p[1]), vsub(t.p[0], t.p[2])));
        t.maxz = std::fmax(std::fmax(q[2], q[5]), q[8]);
    }
    const std::vector<uint32_t> fid = feature_ids(tris, ntri, order.data());
    const size_t nbox = (ntri + kLeafH - 1) / kLeafH;
    std::vector<BoxOrigin> org(nbox);
    std::vector<double> bbx(4 * nbox);
    for (size_t b = 0; b < nbox; ++b) {
        double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY, z0 = INFINITY, z1 = -INFINITY;
        for (size_t k = b * kLeafH; k < std::min(ntri, (b + 1) * kLeafH); ++k)
            for (int j = 0; j < 3; ++j) {
                x0 = std::fmin(x0, tx[k].p[j].x); x1 = std::fmax(x1, tx[k].p[j].x);
                y0 = std::fmin(y0, tx[k].p[j].y); y1 = std::fmax(y1, tx[k].p[j].y);
                z0 = std::fmin(z0, tx[k].p[j].z); z1 = std::fmax(z1, tx[k].p[j].z);
            }
        // same rounding as the device boxes (float, outward)
        const float fx0 = std::nextafterf((float)x0, -INFINITY), fx1 = std::nextafterf((float)x1, INFINITY);
        const float fy0 = std::nextafterf((float)y0, -INFINITY), fy1 = std::nextafterf((float)y1, INFINITY);
        bbx[4 * b] = fx0; bbx[4 * b + 1] = fx1; bbx[4 * b + 2] = fy0; bbx[4 * b + 3] = fy1;
        org[b].ox = 0.5 * ((double)fx0 + (double)fx1);
        org[b].oy = 0.5 * ((double)fy0 + (double)fy1);
        org[b].oz = 0.5 * (z0 + z1);
        org[b].hx = std::nextafterf((float)(0.5 * ((double)fx1 - (double)fx0)), INFINITY);
        org[b].hy = std::nextafterf((float)(0.5 * ((double)fy1 - (double)fy0)), INFINITY);
    }
    const bool comp = cd.c.kind == K_COMPOSITE;
    if (comp && (!comp_fast(cd) || !use_float)) return OCL_E_UNSUPPORTED;
    auto go = [&](auto tag) {
        typedef decltype(tag) T;
        SelErr er;
        er.u = FM<T>::eps() * escale;
        er.R = cd.c.radius;
        er.zmag = 4.0 * cd.c.radius + 10.0;
        SelParams<T> PS[2];
        SelParams<T> P = comp ? comp_sel_params<T>(cd, er, PS) : params<T>(cd, er);
        const size_t stride = (comp ? 2 : 1) * (size_t)rec_len<T>();
        std::vector<T> rec(stride * ntri);
        std::vector<T> boxcap(nbox, -FM<T>::inf());
        double emax = 0;
        for (size_t k = 0; k < ntri; ++k) {
            const BoxOrigin& o = org[k / kLeafH];
            T* r = &rec[stride * k];
            if (comp) {
                build_record_comp<T>(tx[k], cd, o.ox, o.oy, o.oz, er, r, &fid[6 * k]);
                const T* r1 = r + rec_len<T>();
                boxcap[k / kLeafH] = std::max(boxcap[k / kLeafH], std::max(r[R_ZCAP], r1[R_ZCAP]));
                emax = std::max(emax, std::max((double)r[R_EMAX], (double)r1[R_EMAX]));
            } else {
                build_record<T>(tx[k], cd, o.ox, o.oy, o.oz, er, r, &fid[6 * k]);
                boxcap[k / kLeafH] = std::max(boxcap[k / kLeafH], r[R_ZCAP]);
                emax = std::max(emax, (double)r[R_EMAX]);
            }
        }
        P.CUT = (T)(2.0 * emax + (double)P.HPAD);
        switch (cd.c.kind) {
        case K_CYL: run_points<T, K_CYL>(tx, org, bbx, cd, P, PS, cl, n, rec, boxcap, order.data(), z_out, type_out, esc_out, feat_out, tri_out); break;
        case K_BALL: run_points<T, K_BALL>(tx, org, bbx, cd, P, PS, cl, n, rec, boxcap, order.data(), z_out, type_out, esc_out, feat_out, tri_out); break;
        case K_BULL: run_points<T, K_BULL>(tx, org, bbx, cd, P, PS, cl, n, rec, boxcap, order.data(), z_out, type_out, esc_out, feat_out, tri_out); break;
        case K_COMPOSITE: run_points<T, K_COMPOSITE>(tx, org, bbx, cd, P, PS, cl, n, rec, boxcap, order.data(), z_out, type_out, esc_out, feat_out, tri_out); break;
        default: run_points<T, K_CONE>(tx, org, bbx, cd, P, PS, cl, n, rec, boxcap, order.data(), z_out, type_out, esc_out, feat_out, tri_out); break;
        }
    };
    if (use_float) go(0.0f);
    else go(0.0);
    return OCL_OK;
}

// Error calibration: for every (point, triangle) pair evaluates each contact
// (facet, 3 vertices, 3 edges) from the float and the double record and
// returns, per contact class, the largest |z_float - z_double| / bound seen
// on valid, well-conditioned contacts (bound = the error bound the selection
// uses for that contact).  ratio[0] facet, [1] vertex, [2] edge.
extern "C" int ocl_selcheck_calibrate(const double* tris, size_t ntri, const ocl_cutter_desc* desc, const double* cl,
                                      size_t n, double* ratio, double* worst_abs) {
    CutterDev cd;
    if (!cutter_of(desc, cd)) return OCL_E_UNSUPPORTED;
    if (cd.c.kind != K_CYL && cd.c.kind != K_BALL) return OCL_E_UNSUPPORTED;
    SelErr e32, e64;
    e32.u = FM<float>::eps(); e32.R = cd.c.radius; e32.zmag = 4.0 * cd.c.radius + 10.0;
    e64 = e32; e64.u = FM<double>::eps();
    const SelParams<float> P = sel_params<float>(cd, e32);
    std::vector<TriX> tx(ntri);
    for (size_t r = 0; r < ntri; ++r) {
        const double* q = tris + 9 * r;
        tx[r].p[0] = v3(q[0], q[1], q[2]); tx[r].p[1] = v3(q[3], q[4], q[5]); tx[r].p[2] = v3(q[6], q[7], q[8]);
        tx[r].n = vunit(vcross(vsub(tx[r].p[0], tx[r].p[1]), vsub(tx[r].p[0], tx[r].p[2])));
        tx[r].maxz = std::fmax(std::fmax(q[2], q[5]), q[8]);
    }
    for (int c = 0; c < 3; ++c) { ratio[c] = 0; worst_abs[c] = 0; }
    const double R = cd.c.radius;
    for (size_t i = 0; i < n; ++i) {
        const double x = cl[3 * i], y = cl[3 * i + 1];
        for (size_t k = 0; k < ntri; ++k) {
            const TriX& t = tx[k];
            double bx0 = std::fmin(std::fmin(t.p[0].x, t.p[1].x), t.p[2].x), bx1 = std::fmax(std::fmax(t.p[0].x, t.p[1].x), t.p[2].x);
            double by0 = std::fmin(std::fmin(t.p[0].y, t.p[1].y), t.p[2].y), by1 = std::fmax(std::fmax(t.p[0].y, t.p[1].y), t.p[2].y);
            if (bx1 < x - R || bx0 > x + R || by1 < y - R || by0 > y + R) continue;
            // box origin: snap to a 2R grid like a level-0 box centre would be
            const double ox = std::floor(bx0 / 2.0) * 2.0 + 1.0, oy = std::floor(by0 / 2.0) * 2.0 + 1.0, oz = t.maxz;
            float rf[kFastRec32];
            double rd[kFastRec32];
            build_record<float>(t, cd, ox, oy, oz, e32, rf);
            build_record<double>(t, cd, ox, oy, oz, e64, rd);
            const float xf = (float)(x - ox), yf = (float)(y - oy);
            const double xd = x - ox, yd = y - oy;
            // facet
            {
                const float u = fmaf(rf[R_U2], yf, fmaf(rf[R_U1], xf, rf[R_U0]));
                const float v = fmaf(rf[R_V2], yf, fmaf(rf[R_V1], xf, rf[R_V0]));
                const float mf = std::fmin(std::fmin(u, v), 1.0f - u - v);
                const float zf = fmaf(rf[R_GY], yf, fmaf(rf[R_GX], xf, rf[R_F0]));
                const double zd = fma(rd[R_GY], yd, fma(rd[R_GX], xd, rd[R_F0]));
                if (mf > rf[R_EBF]) {
                    const double err = std::fabs((double)zf - zd);
                    ratio[0] = std::fmax(ratio[0], err / (double)rf[R_EZF]);
                    worst_abs[0] = std::fmax(worst_abs[0], err);
                }
            }
            for (int j = 0; j < 3; ++j) {
                const float dx = rf[R_P0X + 3 * j] - xf, dy = rf[R_P0Y + 3 * j] - yf;
                const float m = P.R2 - fmaf(dx, dx, dy * dy);
                const double ddx = rd[R_P0X + 3 * j] - xd, ddy = rd[R_P0Y + 3 * j] - yd;
                const double md = R * R - fma(ddx, ddx, ddy * ddy);
                if (m > P.S2V) {
                    const float zf = rf[R_P0Z + 3 * j] + (cd.c.kind == K_BALL ? std::sqrt(m) : 0.0f);
                    const double zd = rd[R_P0Z + 3 * j] + (cd.c.kind == K_BALL ? std::sqrt(md) : 0.0);
                    const double err = std::fabs((double)zf - zd);
                    ratio[1] = std::fmax(ratio[1], err / (double)P.E);
                    worst_abs[1] = std::fmax(worst_abs[1], err);
                }
                const float* ef = rf + R_E0 + kEdge * j;
                const double* ed = rd + R_E0 + kEdge * j;
                const float sd = fmaf(ef[E_EX], yf, fmaf(-ef[E_EY], xf, ef[E_C]));
                const float me = fmaf(-sd, sd, P.R2);
                const float u1 = fmaf(-ef[E_EY], yf, fmaf(-ef[E_EX], xf, ef[E_A]));
                const double sdd = fma(ed[E_EX], yd, fma(-ed[E_EY], xd, ed[E_C]));
                const double med = fma(-sdd, sdd, R * R);
                const double u1d = fma(-ed[E_EY], yd, fma(-ed[E_EX], xd, ed[E_A]));
                if (me > ef[E_S2E] && med > 0) {
                    const float s = std::sqrt(me);
                    const double sdb = std::sqrt(med);
                    const float zf = fmaf(s, ef[E_K1], fmaf(-u1, ef[E_M], ef[E_Z1]));
                    const double zd = fma(sdb, ed[E_K1], fma(-u1d, ed[E_M], ed[E_Z1]));
                    const double err = std::fabs((double)zf - zd);
                    ratio[2] = std::fmax(ratio[2], err / (double)ef[E_EZE]);
                    worst_abs[2] = std::fmax(worst_abs[2], err);
                }
            }
        }
    }
    return OCL_OK;
}

// The push kernel's step-wise Brent (brent_step) against aell_brent on the
// BullCutter edge ellipses of n (fiber, edge) cases.  Input per case (14
// doubles): fiber p1, p2; edge p1, p2; diameter, corner radius.  Output per
// case (9 doubles): needs-solve flag, then for each root (d1, d2): the root
// and step count of aell_brent, then of the state machine.
extern "C" int ocl_selcheck_brent(const double* in, size_t n, double* out) {
    for (size_t i = 0; i < n; ++i) {
        const double* q = in + 14 * i;
        double* o = out + 9 * i;
        for (int k = 0; k < 9; ++k) o[k] = 0.0;
        const CutterP c = CutterP::bull(q[12], q[13], 20.0);
        const FibX f = make_fiber(q);
        const V3 p1 = v3(q[6], q[7], q[8]), p2 = v3(q[9], q[10], q[11]);
        AEll e;
        double lo, hi;
        if (!bull_edge_prep(c, f, p1, p2, e) || !bull_edge_bracket(e, f, lo, hi)) continue;
        o[0] = 1.0;
        for (int which = 0; which < 2; ++which) {
            int it = 0;
            const double r = bull_edge_solve(e, lo, hi, which, it);
            BrentState B;
            if (which == 0) brent_start(B, lo, hi);
            else brent_start(B, hi - 4.0, lo);
            while (!brent_step(B, aell_error(e, brent_x(B)), 3E-16, 1e-10)) {
            }
            o[1 + 4 * which] = r;
            o[2 + 4 * which] = it;
            o[3 + 4 * which] = B.b;
            o[4 + 4 * which] = B.it;
        }
    }
    return 0;
}


// The selection's triangle cap (tri_cap: profile cap and plane cap) against
// the triangle's own contacts.  For every (point, triangle) pair of the
// inputs, with the triangle's record built in a frame at its centroid:
// cap_out = the cap at the point and hi_out = the largest upper value
// (z + error bound, or z + slack) of the contacts the selection evaluates
// from z = -inf (-inf: none), both in the global frame; n x ntri row-major.
extern "C" int ocl_selcheck_tricap(const double* tris, size_t ntri, const ocl_cutter_desc* desc, const double* pts,
                                   size_t n, int use_float, double* cap_out, double* hi_out) {
    CutterDev cd;
    if (!cutter_of(desc, cd)) return OCL_E_UNSUPPORTED;
    const bool comp = cd.c.kind == K_COMPOSITE;
    if (comp && (!comp_fast(cd) || !use_float)) return OCL_E_UNSUPPORTED;
    auto go = [&](auto tag) {
        typedef decltype(tag) T;
        typedef FM<T> F;
        SelErr er;
        er.u = F::eps();
        er.R = cd.c.radius;
        er.zmag = 4.0 * cd.c.radius + 10.0;
        SelParams<T> PS[2];
        const SelParams<T> P = comp ? comp_sel_params<T>(cd, er, PS) : params<T>(cd, er);
        T r[2 * kFastRec32];
        for (size_t k = 0; k < ntri; ++k) {
            const double* q = tris + 9 * k;
            TriX t;
            t.p[0] = v3(q[0], q[1], q[2]);
            t.p[1] = v3(q[3], q[4], q[5]);
            t.p[2] = v3(q[6], q[7], q[8]);
            t.n = vunit(vcross(vsub(t.p[0], t.p[1]), vsub(t.p[0], t.p[2])));
            t.maxz = std::fmax(std::fmax(q[2], q[5]), q[8]);
            const double ox = (q[0] + q[3] + q[6]) / 3.0, oy = (q[1] + q[4] + q[7]) / 3.0;
            const double oz = 0.5 * (std::fmin(std::fmin(q[2], q[5]), q[8]) + t.maxz);
            if (comp) build_record_comp<T>(t, cd, ox, oy, oz, er, r, nullptr);
            else build_record<T>(t, cd, ox, oy, oz, er, r);
            for (size_t i = 0; i < n; ++i) {
                const T xl = (T)(pts[3 * i] - ox), yl = (T)(pts[3 * i + 1] - oy);
                const T d2 = box_d2<T>(xl, yl, r[R_BX0], r[R_BX1], r[R_BY0], r[R_BY1]);
                if (comp) {
                    // the lower of the triangle cap and the box-level bound
                    // (the largest record cap less the profile's lower bound)
                    const T* r1 = r + rec_len<T>();
                    const T cb = F::max_(r[R_ZCAP], r1[R_ZCAP]) - comp_prof_lower<T>(d2, P);
                    const T cap = F::min_(comp_tri_cap<T>(r, xl, yl, d2, P), cb);
                    T hmax = -F::inf();
                    const T zf0 = F::fma_(r[R_GY], yl, F::fma_(r[R_GX], xl, r[R_F0]));
                    const T zf1 = F::fma_(r1[R_GY], yl, F::fma_(r1[R_GX], xl, r1[R_F0]));
                    const T starts[4] = {-F::inf(), zf0, zf1, F::max_(zf0, zf1)};
                    for (T z0 : starts) {
                        Sel<T> S;
                        sel_init(S, z0);
                        T hi = -F::inf();
                        sel_triangle_comp<T>(r, 0, xl, yl, P, PS, S, true, &hi);
                        hmax = F::max_(hmax, hi);
                    }
                    cap_out[i * ntri + k] = (double)cap + oz;
                    hi_out[i * ntri + k] = (double)hmax + oz;
                    continue;
                }
                T cap = 0;
                switch (cd.c.kind) {
                case K_CYL: cap = tri_cap<T, K_CYL>(r, xl, yl, d2, P); break;
                case K_BALL: cap = tri_cap<T, K_BALL>(r, xl, yl, d2, P); break;
                case K_BULL: cap = tri_cap<T, K_BULL>(r, xl, yl, d2, P); break;
                default: cap = tri_cap<T, K_CONE>(r, xl, yl, d2, P); break;
                }
                // from -inf (a valid facet contact lifts and ends the
                // triangle), then from the facet height itself (the facet
                // does not lift, the vertices and edges run)
                const T zf = F::fma_(r[R_GY], yl, F::fma_(r[R_GX], xl, r[R_F0]));
                T hmax = -F::inf();
                for (int pass = 0; pass < 2; ++pass) {
                    Sel<T> S;
                    sel_init(S, pass == 0 ? -F::inf() : zf);
                    T hi = -F::inf();
                    switch (cd.c.kind) {
                    case K_CYL: sel_triangle<T, K_CYL>(r, 0, xl, yl, P, S, true, &hi); break;
                    case K_BALL: sel_triangle<T, K_BALL>(r, 0, xl, yl, P, S, true, &hi); break;
                    case K_BULL: sel_triangle<T, K_BULL>(r, 0, xl, yl, P, S, true, &hi); break;
                    default: sel_triangle<T, K_CONE>(r, 0, xl, yl, P, S, true, &hi); break;
                    }
                    hmax = F::max_(hmax, hi);
                }
                cap_out[i * ntri + k] = (double)cap + oz;
                hi_out[i * ntri + k] = (double)hmax + oz;
            }
        }
    };
    if (use_float) go(0.0f);
    else go(0.0);
    return OCL_OK;
}

// One (point, triangle) pair through sel_triangle from an empty lane, the
// record built in the frame of the triangle's own box centre.  out (12):
// zb, eb, feature, zt[0..3], za, zr, live pairs, facet (u,v) margin, unused.
extern "C" int ocl_selcheck_pair(const double* tri9, const ocl_cutter_desc* desc, double x, double y, int use_float,
                                 double* out) {
    CutterDev cd;
    if (!cutter_of(desc, cd)) return OCL_E_UNSUPPORTED;
    TriX t;
    t.p[0] = v3(tri9[0], tri9[1], tri9[2]);
    t.p[1] = v3(tri9[3], tri9[4], tri9[5]);
    t.p[2] = v3(tri9[6], tri9[7], tri9[8]);
    t.n = vunit(vcross(vsub(t.p[0], t.p[1]), vsub(t.p[0], t.p[2])));
    t.maxz = std::fmax(std::fmax(tri9[2], tri9[5]), tri9[8]);
    const double ox = 0.5 * (std::fmin(std::fmin(tri9[0], tri9[3]), tri9[6]) + std::fmax(std::fmax(tri9[0], tri9[3]), tri9[6]));
    const double oy = 0.5 * (std::fmin(std::fmin(tri9[1], tri9[4]), tri9[7]) + std::fmax(std::fmax(tri9[1], tri9[4]), tri9[7]));
    const double oz = t.maxz;
    auto go = [&](auto tag) {
        typedef decltype(tag) T;
        SelErr er;
        er.u = FM<T>::eps();
        er.R = cd.c.radius;
        er.zmag = 4.0 * cd.c.radius + 10.0;
        SelParams<T> P = params<T>(cd, er);
        T rec[kFastRec32];
        build_record<T>(t, cd, ox, oy, oz, er, rec);
        P.CUT = (T)(2.0 * (double)rec[R_EMAX] + (double)P.HPAD);
        Sel<T> S;
        sel_init(S, -FM<T>::inf());
        const T xl = (T)(x - ox), yl = (T)(y - oy);
        switch (cd.c.kind) {
        case K_CYL: sel_triangle<T, K_CYL>(rec, 0, xl, yl, P, S, true); break;
        case K_BALL: sel_triangle<T, K_BALL>(rec, 0, xl, yl, P, S, true); break;
        case K_BULL: sel_triangle<T, K_BULL>(rec, 0, xl, yl, P, S, true); break;
        default: sel_triangle<T, K_CONE>(rec, 0, xl, yl, P, S, true); break;
        }
        out[0] = oz + (double)S.zb;
        out[1] = S.eb;
        out[2] = S.bf;
        for (int q = 0; q < 4; ++q) out[3 + q] = oz + (double)S.zt[q];
        out[7] = oz + (double)S.za;
        out[8] = oz + (double)S.zr;
        out[9] = S.pairs;
        out[10] = (double)rec[R_EMAX];
        out[11] = (double)rec[R_FLAGS];
    };
    if (use_float) go(0.0f);
    else go(0.0);
    return OCL_OK;
}

// Feature IDs of a triangle soup in input order (6 per triangle: vertices
// p0..p2, edges (p0,p1), (p1,p2), (p2,p0)), as the mesh upload computes them.
extern "C" int ocl_selcheck_feature_ids(const double* tris, size_t ntri, uint32_t* out) {
    const std::vector<uint32_t> ids = feature_ids(tris, ntri, nullptr);
    std::memcpy(out, ids.data(), sizeof(uint32_t) * ids.size());
    return OCL_OK;
}

// The push kernel's pair bound (push_bound_rec + push_bound) for one
// (fiber, triangle) pair: returns 0 (no bound), 1 (no interval) or 2 (any
// interval in [out[0], out[1]]); dir 0 = X fiber, 1 = Y fiber.
extern "C" int ocl_selcheck_push_bound(const double* tri9, const ocl_cutter_desc* desc, const double* fib6, int dir,
                                       double* out) {
    CutterDev cd;
    if (!cutter_of(desc, cd)) return -1;
    double rec[3];
    push_bound_rec(tri9, dir == 0 ? 0 : 1, rec);
    const double zmax = std::max(std::max(tri9[2], tri9[5]), tri9[8]);
    const int ca = dir == 0 ? 1 : 0;
    const double c0 = std::min(std::min(tri9[ca], tri9[3 + ca]), tri9[6 + ca]);
    const double c1 = std::max(std::max(tri9[ca], tri9[3 + ca]), tri9[6 + ca]);
    double tl = 0.0, tu = 0.0;
    const int r = push_bound(cd.c, rec, zmax, c0, c1, fib6[2], fib6[ca], fib6[dir == 0 ? 0 : 1],
                             fib6[dir == 0 ? 3 : 4], tl, tu);
    out[0] = tl;
    out[1] = tu;
    return r;
}
