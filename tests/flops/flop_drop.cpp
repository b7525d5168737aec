// Flops per evaluated (CL point, triangle) pair of the FP32 selection's
// contact function (sel_triangle in drop_fast.hpp), counted by running that
// function on the host with a counting 4-byte float type.  Counting rule:
// add/sub/mul = 1, fma = 2, sqrt/rsqrt/rcp/div = 1; compare, select,
// min/max, abs and negation = 0.  sel_triangle has no data-dependent
// branches (every contact runs, validity only masks its result), so one pair
// gives the count of every pair.  Test infrastructure only.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace fcount {
struct Tally {
    unsigned long long add, mul, fma, sqrt, div;
};
static Tally g;

struct CF {
    float v;
    CF() = default;
    constexpr CF(float a) : v(a) {}
    constexpr CF(double a) : v((float)a) {}
    constexpr CF(int a) : v((float)a) {}
    constexpr CF(unsigned a) : v((float)a) {}
    explicit constexpr operator float() const { return v; }
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator int() const { return (int)v; }
    explicit constexpr operator unsigned() const { return (unsigned)v; }
    CF& operator+=(CF b) { ++g.add; v += b.v; return *this; }
    CF& operator-=(CF b) { ++g.add; v -= b.v; return *this; }
    CF& operator*=(CF b) { ++g.mul; v *= b.v; return *this; }
};
inline CF operator+(CF a, CF b) { ++g.add; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { ++g.add; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { ++g.mul; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { ++g.div; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
inline bool operator==(CF a, CF b) { return a.v == b.v; }
inline bool operator!=(CF a, CF b) { return a.v != b.v; }
inline CF fabs(CF a) { return CF(std::fabs(a.v)); }
static_assert(sizeof(CF) == 4, "the selection takes its FP32 branches by sizeof(T)");
}  // namespace fcount

#include "../../opencamlib_amd/csrc/drop_fast.hpp"

namespace ocl_amd {
template <> struct FM<fcount::CF> {
    typedef fcount::CF CF;
    static CF sqrt_(CF a) { ++fcount::g.sqrt; return CF(std::sqrt(a.v)); }
    static CF rsqrt_(CF a) { ++fcount::g.sqrt; return CF(1.0f / std::sqrt(a.v)); }
    static CF rcp_(CF a) { ++fcount::g.div; return CF(1.0f / a.v); }
    static CF fma_(CF a, CF b, CF c) { ++fcount::g.fma; return CF(std::fma(a.v, b.v, c.v)); }
    static CF max_(CF a, CF b) { return CF(std::fmax(a.v, b.v)); }
    static CF min_(CF a, CF b) { return CF(std::fmin(a.v, b.v)); }
    static constexpr CF inf() { return CF((float)INFINITY); }
    static constexpr double eps() { return FM<float>::eps(); }
};
}  // namespace ocl_amd

using namespace ocl_amd;

namespace {
template <int K>
void one_pair(const CutterDev& cd, const TriX& t, double x, double y) {
    typedef fcount::CF CF;
    SelErr er;
    er.u = FM<CF>::eps();
    er.R = cd.c.radius;
    er.zmag = 4.0 * cd.c.radius + 10.0;
    SelParams<CF> P = sel_params<CF>(cd, er);
    CF rec[kFastRec32];
    const uint32_t fid[6] = {0, 1, 2, 3, 4, 5};
    build_record<CF>(t, cd, 0.0, 0.0, 0.0, er, rec, fid);
    P.CUT = CF(1.0f);
    Sel<CF> S;
    sel_init(S, -FM<CF>::inf());
    std::memset(&fcount::g, 0, sizeof(fcount::g));
    sel_triangle<CF, K>(rec, 0, CF(x), CF(y), P, S, true);
}
}  // namespace

// kind: 0 cyl, 1 ball, 2 bull, 3 cone (d = 10, bull r = 2, cone 45 deg);
// out: add, mul, fma, sqrt (incl. rsqrt), div (incl. rcp) of one pair.
extern "C" __attribute__((visibility("default"))) int fc_drop_pair(int kind, unsigned long long* out) {
    CutterDev cd;
    std::memset(&cd, 0, sizeof(cd));
    switch (kind) {
    case 0: cd.c = CutterP::cyl(10.0, 20.0); break;
    case 1: cd.c = CutterP::ball(10.0, 20.0); break;
    case 2: cd.c = CutterP::bull(10.0, 2.0, 20.0); break;
    case 3: cd.c = CutterP::cone(10.0, M_PI / 4, 10.0); break;
    default: return 1;
    }
    TriX t;
    t.p[0] = v3(0.0, 0.0, 1.0);
    t.p[1] = v3(3.0, 0.5, 1.5);
    t.p[2] = v3(1.0, 2.5, 2.0);
    t.n = vunit(vcross(vsub(t.p[0], t.p[1]), vsub(t.p[0], t.p[2])));
    t.maxz = 2.0;
    switch (kind) {
    case 0: one_pair<K_CYL>(cd, t, 1.2, 1.0); break;
    case 1: one_pair<K_BALL>(cd, t, 1.2, 1.0); break;
    case 2: one_pair<K_BULL>(cd, t, 1.2, 1.0); break;
    default: one_pair<K_CONE>(cd, t, 1.2, 1.0); break;
    }
    out[0] = fcount::g.add;
    out[1] = fcount::g.mul;
    out[2] = fcount::g.fma;
    out[3] = fcount::g.sqrt;
    out[4] = fcount::g.div;
    return 0;
}
