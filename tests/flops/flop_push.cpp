// Flops per (fiber, triangle) pair of the FP64 push contact code
// (push_pair in push_exact.hpp, what k_push_fibers runs per candidate pair),
// counted by compiling that code on the host with `double` replaced by a
// counting type and running it over given pairs.  Counting rule: add/sub/
// mul = 1, fma = 2, sqrt/div = 1, every other libm call (atan2, sin, cos,
// ...) = 1; compare, select, min/max, abs and negation = 0.  The push code
// branches on the geometry, so the count is an average over the pairs given
// (the caller passes a sample of the workload's own candidate pairs).  Brent
// iterations of the BullCutter edge solves are counted on their own, with
// the flops they spend.  Test infrastructure only.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace fcount {
struct Tally {
    unsigned long long add, mul, fma, sqrt, div, other;
};
static Tally g;

struct CD {
    double v;
    CD() = default;
    constexpr CD(double a) : v(a) {}
    constexpr CD(float a) : v(a) {}
    constexpr CD(int a) : v(a) {}
    constexpr CD(unsigned a) : v(a) {}
    constexpr CD(long a) : v((double)a) {}
    constexpr CD(unsigned long a) : v((double)a) {}
    constexpr CD(long long a) : v((double)a) {}
    constexpr CD(unsigned long long a) : v((double)a) {}
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator float() const { return (float)v; }
    explicit constexpr operator int() const { return (int)v; }
    explicit constexpr operator unsigned() const { return (unsigned)v; }
    explicit constexpr operator bool() const { return v != 0.0; }
    CD& operator+=(CD b) { ++g.add; v += b.v; return *this; }
    CD& operator-=(CD b) { ++g.add; v -= b.v; return *this; }
    CD& operator*=(CD b) { ++g.mul; v *= b.v; return *this; }
    CD& operator/=(CD b) { ++g.div; v /= b.v; return *this; }
};
inline CD operator+(CD a, CD b) { ++g.add; return CD(a.v + b.v); }
inline CD operator-(CD a, CD b) { ++g.add; return CD(a.v - b.v); }
inline CD operator*(CD a, CD b) { ++g.mul; return CD(a.v * b.v); }
inline CD operator/(CD a, CD b) { ++g.div; return CD(a.v / b.v); }
inline CD operator-(CD a) { return CD(-a.v); }
inline CD operator+(CD a) { return a; }
inline bool operator<(CD a, CD b) { return a.v < b.v; }
inline bool operator>(CD a, CD b) { return a.v > b.v; }
inline bool operator<=(CD a, CD b) { return a.v <= b.v; }
inline bool operator>=(CD a, CD b) { return a.v >= b.v; }
inline bool operator==(CD a, CD b) { return a.v == b.v; }
inline bool operator!=(CD a, CD b) { return a.v != b.v; }
inline CD fabs(CD a) { return CD(std::fabs(a.v)); }
inline CD fmin(CD a, CD b) { return CD(std::fmin(a.v, b.v)); }
inline CD fmax(CD a, CD b) { return CD(std::fmax(a.v, b.v)); }
inline CD sqrt(CD a) { ++g.sqrt; return CD(std::sqrt(a.v)); }
inline CD fma(CD a, CD b, CD c) { ++g.fma; return CD(std::fma(a.v, b.v, c.v)); }
inline CD copysign(CD a, CD b) { return CD(std::copysign(a.v, b.v)); }
inline CD floor(CD a) { return CD(std::floor(a.v)); }
inline CD atan2(CD a, CD b) { ++g.other; return CD(std::atan2(a.v, b.v)); }
inline CD atan(CD a) { ++g.other; return CD(std::atan(a.v)); }
inline CD sin(CD a) { ++g.other; return CD(std::sin(a.v)); }
inline CD cos(CD a) { ++g.other; return CD(std::cos(a.v)); }
inline CD tan(CD a) { ++g.other; return CD(std::tan(a.v)); }
inline bool isnan(CD a) { return std::isnan(a.v); }
inline bool isfinite(CD a) { return std::isfinite(a.v); }
inline bool isinf(CD a) { return std::isinf(a.v); }
inline bool signbit(CD a) { return std::signbit(a.v); }
}  // namespace fcount
// cutter.hpp calls std::tan by name (a set-up cost, counted like the rest)
namespace std {
inline fcount::CD tan(fcount::CD a) { return fcount::tan(a); }
}  // namespace std

// the push code in a namespace of its own, with `double` counted
#define ocl_amd ocl_amd_counted
#define double fcount::CD
#include "../../opencamlib_amd/csrc/push_exact.hpp"
#undef double

using namespace ocl_amd_counted;

namespace {
void snapshot(unsigned long long* out) {
    out[0] = fcount::g.add;
    out[1] = fcount::g.mul;
    out[2] = fcount::g.fma;
    out[3] = fcount::g.sqrt;
    out[4] = fcount::g.div;
    out[5] = fcount::g.other;
}
}  // namespace

// kind 2: BullCutter d r L (spec[0..2]); kind 3: ConeCutter d angle L.
// tris: ntri x 9 doubles; fib: nf x 6; pairs: npairs (fiber, triangle)
// index pairs.  out[0..5]: add, mul, fma, sqrt, div, other over all pairs;
// out[6]: Brent iterations (their flops are part of out[0..5]; fc_brent_solve
// measures one iteration's cost).
extern "C" __attribute__((visibility("default"))) int fc_push_pairs(int kind, const double* spec, const double* tris,
                                                                    const double* fib, const uint32_t* pairs,
                                                                    size_t npairs, unsigned long long* out) {
    CutterDev cd;
    std::memset((void*)&cd, 0, sizeof(cd));
    if (kind == 2) cd.c = CutterP::bull(spec[0], spec[1], spec[2]);
    else if (kind == 3) cd.c = CutterP::cone(spec[0], spec[1], spec[2]);
    else return 1;
    std::memset(&fcount::g, 0, sizeof(fcount::g));
    unsigned long long brent_it = 0;
    for (size_t k = 0; k < npairs; ++k) {
        const double* f = fib + 6 * (size_t)pairs[2 * k];
        const double* q = tris + 9 * (size_t)pairs[2 * k + 1];
        fcount::CD fq[6];
        for (int j = 0; j < 6; ++j) fq[j] = fcount::CD(f[j]);
        const fcount::Tally before_setup = fcount::g;
        const FibX fx = make_fiber(fq);
        TriX t;
        t.p[0] = v3(q[0], q[1], q[2]);
        t.p[1] = v3(q[3], q[4], q[5]);
        t.p[2] = v3(q[6], q[7], q[8]);
        t.n = vunit(vcross(vsub(t.p[0], t.p[1]), vsub(t.p[0], t.p[2])));
        t.maxz = std::max(std::max(q[2], q[5]), q[8]);
        fcount::g = before_setup;  // fiber and normal set-up are per-fiber / per-mesh work
        PIval iv;
        piv_init(iv);
        int it = 0;
        if (kind == 2) push_pair<K_BULL>(cd, fx, t, iv, it);
        else push_pair<K_CONE>(cd, fx, t, iv, it);
        brent_it += (unsigned long long)it;
    }
    snapshot(out);
    out[6] = brent_it;
    return 0;
}

// Flops of Brent iterations: runs one BullCutter edge solve (the first pair
// that needs one) with the iteration counter, and reports its flops and
// iterations (out[0..5], out[6]).
extern "C" __attribute__((visibility("default"))) int fc_brent_solve(const double* spec, const double* tris,
                                                                     const double* fib, const uint32_t* pairs,
                                                                     size_t npairs, unsigned long long* out) {
    CutterDev cd;
    std::memset((void*)&cd, 0, sizeof(cd));
    cd.c = CutterP::bull(spec[0], spec[1], spec[2]);
    for (size_t k = 0; k < npairs; ++k) {
        const double* f = fib + 6 * (size_t)pairs[2 * k];
        const double* q = tris + 9 * (size_t)pairs[2 * k + 1];
        fcount::CD fq[6];
        for (int j = 0; j < 6; ++j) fq[j] = fcount::CD(f[j]);
        const FibX fx = make_fiber(fq);
        for (int e = 0; e < 3; ++e) {
            const V3 p1 = v3(q[3 * e], q[3 * e + 1], q[3 * e + 2]);
            const int e2 = (e + 1) % 3;
            const V3 p2 = v3(q[3 * e2], q[3 * e2 + 1], q[3 * e2 + 2]);
            AEll el;
            fcount::CD lo = 0.0, hi = 0.0;
            CutterP c = cd.c;
            if (!bull_edge_prep(c, fx, p1, p2, el) || !bull_edge_bracket(el, fx, lo, hi)) continue;
            std::memset(&fcount::g, 0, sizeof(fcount::g));
            int it = 0;
            bull_edge_solve(el, lo, hi, 0, it);
            snapshot(out);
            out[6] = (unsigned long long)it;
            return 0;
        }
    }
    return 1;
}
