// Batch push-cutter kernels (BatchPushCutter::pushCutter3,
// src/algo/batchpushcutter.cpp:136-188).
//
// The mesh copy used here is ordered by the reference's YZ (X fibers) or XZ
// (Y fibers) KD-tree rank, with box hierarchies over (across, z) and over the
// KD search's acceptance bounds.  Each fiber's interval list
// (Fiber::addInterval, fiber.cpp:58-93) is kept sorted and disjoint.
#pragma once

#include "mesh.hpp"
#include "push_exact.hpp"
#include "walk.hpp"

namespace ocl_amd {

// One union interval: lower/upper fiber parameter with CC points and types.
struct PEntry {
    double lo, hi;
    double lcc[3], ucc[3];
    int lty, uty;
};

// Where the fibers' lists go.  First pass (rcap == 0): a list is built in
// LDS (lds_list<K>() entries) and copied into a compact pool at an offset taken
// from a counter; a fiber whose list outgrows LDS, or does not fit in the
// pool, is appended to `retry` (and written with 0 entries).  Retry pass
// (rcap > 0, over the `retry` fibers only): slot s of the pass builds its list
// in global memory at big + s * rcap; a list that outgrows rcap counts in
// *retry_n and the pass runs again with a larger rcap.  Intermediate lists
// can be longer than the final one (intervals met early merge later), so a
// list's final length does not bound what it needs while it is built.
constexpr uint64_t kBigList = 1ull << 63;  // off[]: the list is in `big`
struct PushLists {
    PEntry* pool;
    unsigned long long* used;  // pool entries taken
    unsigned long long pool_cap;
    uint64_t* off;             // per fiber (input index): offset in pool, or kBigList | offset in big
    uint32_t* n;               // entries per fiber
    uint32_t* retry;           // first pass: fibers to retry (output); retry pass: the fibers (input)
    uint32_t* retry_n;         // first pass: their count; retry pass: lists that outgrew rcap
    PEntry* big;
    uint32_t rcap;
};

// One fiber direction of a push launch: the direction's tree (rank order
// and boxes), the KD acceptance bounds, its fibers in the order they are
// taken (perm: sorted by z, then across) and its interval lists.
struct PushSide {
    MeshView pm;
    const double* acc;
    const float4* reach;  // reach boxes (pair bound), same layout as pm.tree.box
    const double* fib;
    const uint32_t* perm;
    uint32_t nf;
    int dir;  // 0: X fibers (across = y), 1: Y fibers (across = x)
    PushLists PL;
};

OCL_HD PEntry pentry_of(const PIval& i) {
    PEntry p;
    p.lo = i.lo;
    p.hi = i.hi;
    p.lcc[0] = i.lcc.x; p.lcc[1] = i.lcc.y; p.lcc[2] = i.lcc.z;
    p.ucc[0] = i.ucc.x; p.ucc[1] = i.ucc.y; p.ucc[2] = i.ucc.z;
    p.lty = i.lty;
    p.uty = i.uty;
    return p;
}

// Fiber::addInterval on a list sorted by lower end with pairwise disjoint
// entries.  Entries touching or overlapping the new interval merge into one
// (strict Interval::outside, interval.cpp:96-103); the merged lower end
// (upper end) keeps the CC of the existing extreme unless the new interval is
// strictly lower (higher), as the reference's sum interval keeps the first
// extreme it meets, with the existing intervals listed before the new one.
OCL_HD bool plist_insert(PEntry* E, uint32_t& n, uint32_t cap, const PEntry& iv) {
    uint32_t j = 0;
    while (j < n && E[j].hi < iv.lo) ++j;
    if (j == n || E[j].lo > iv.hi) {
        if (n == cap) return false;
        for (uint32_t k = n; k > j; --k) E[k] = E[k - 1];
        E[j] = iv;
        ++n;
        return true;
    }
    // existing entries overlapping iv: E[j..k); being disjoint, they have
    // distinct extremes, and the new interval only wins a strictly lower
    // lower end or a strictly higher upper end
    PEntry m = E[j];
    uint32_t k = j + 1;
    while (k < n && E[k].lo <= iv.hi) {
        m.hi = E[k].hi;
        m.ucc[0] = E[k].ucc[0]; m.ucc[1] = E[k].ucc[1]; m.ucc[2] = E[k].ucc[2];
        m.uty = E[k].uty;
        ++k;
    }
    if (iv.lo < m.lo) {
        m.lo = iv.lo;
        m.lcc[0] = iv.lcc[0]; m.lcc[1] = iv.lcc[1]; m.lcc[2] = iv.lcc[2];
        m.lty = iv.lty;
    }
    if (iv.hi > m.hi) {
        m.hi = iv.hi;
        m.ucc[0] = iv.ucc[0]; m.ucc[1] = iv.ucc[1]; m.ucc[2] = iv.ucc[2];
        m.uty = iv.uty;
    }
    E[j] = m;
    const uint32_t gone = k - (j + 1);
    if (gone) {
        for (uint32_t q = j + 1; q + gone < n; ++q) E[q] = E[q + gone];
        n -= gone;
    }
    return true;
}

struct PushCounters {
    unsigned long long pairs, intervals, brent, skipped;
};

// One wavefront per fiber (persistent grid): the wave walks the tree with
// the fiber's search box, gathers the candidate triangles in rank order in
// batches of 64 (one per lane), computes the 64 pair intervals in parallel
// (FP64) and lane 0 merges them into the fiber's list in rank order.
// K = cutter kind.  One launch serves one or two sides (fiber directions,
// Waterline's X and Y fibers): the sides' fibers form one queue, so the
// second direction fills the GPU while the first one's last fibers finish.
//
// Candidates (superset = 1, the reference's behaviour): every triangle the
// reference's KD search returns for the fiber's search box, reproduced
// exactly from the per-rank leaf acceptance bounds (acc, see
// kd_rank_order); the reference pushes all of them, without an overlap test
// (batchpushcutter.cpp:170-181).  The tree's boxes are then the KD
// acceptance boxes, so the walk's overlap test visits every box holding such
// a triangle.  superset = 0: only triangles whose boxes overlap the search box.
//
// BullCutter: a lane's pair needs up to six Brent solves (two per edge), of
// very different lengths, so a batch runs in three steps: each lane lists the
// solves its pair needs and their brackets; the wavefront runs the listed
// solves 64 at a time, one per lane; each lane then replays its pair in the
// reference's order (vertices, facet, edges) with the solved roots.  A batch
// then costs the longest solve of each round of 64 instead of the longest
// lane's six solves in a row.
//
// skip = 1: a candidate whose interval bound (push_bound) lies strictly
// inside an interval of the fiber's list when it is queued, or whose
// triangle lies below the fiber or beyond the radius across it, is left
// out: it cannot change the list (push_bound's comment), so the results are
// those of pushing every candidate.  The walk first skips whole level-0
// boxes by their reach boxes (k_box_leaf_reach).  The containment test
// reads the list lane 0 keeps in LDS, which only grows and merges while the
// fiber's candidates are taken in rank order; the retry pass, whose list is
// in global memory, only leaves out pairs that cannot reach the fiber.
//
// Merging: a batch's non-empty intervals are first merged among themselves
// in parallel (sort by lower end, components by a prefix max), so lane 0
// inserts a few components rather than every interval; one by one, the
// merges took 65% of the ConeCutter push.  The fiber's list lives in LDS
// while it is built (first pass) and is copied out at the end; the few
// fibers whose lists outgrow it stop there and run again, alone, in a retry
// pass that merges in global memory (PushLists).
constexpr int kPushWaves = 4;  // wavefronts per 256-thread block
constexpr int kBullSolves = 6; // Brent solves per pair: 3 edges x 2 roots
// Interval list entries per fiber in LDS: 128 (61 KB per 4-wavefront block,
// two blocks per CU); the BullCutter's solve tables leave room for 64.
template <int K>
constexpr uint32_t lds_list() {
    return K == K_BULL ? 64u : 128u;
}

template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_push_fibers(const PushSide* __restrict__ sides, uint32_t nsides, uint32_t* __restrict__ queue, int superset,
              int skip, CutterDev cd, PushCounters* cnt) {
    constexpr int kS = (K == K_BULL) ? kBullSolves : 1;
    __shared__ uint32_t qmem[kPushWaves][128];
    __shared__ PEntry emem[kPushWaves][64];
    __shared__ uint32_t tmem[kPushWaves][kS * 64];  // Bull: listed solves (lane << 3 | edge << 1 | root)
    __shared__ double bmem[kPushWaves][kS][64];     // Bull: brackets (lo, hi per edge)
    __shared__ double rmem[kPushWaves][kS][64];     // Bull: roots (d1, d2 per edge)
    constexpr uint32_t kLdsList = lds_list<K>();
    __shared__ PEntry lmem[kPushWaves][kLdsList];   // the fiber's interval list
    const uint32_t lane = lane_id();
    const uint32_t wv = threadIdx.x >> 6;
    uint32_t* qbuf = qmem[wv];
    PEntry* ebuf = emem[wv];
    const double R = cd.c.radius, L = cd.c.length;
    unsigned long long pairs = 0, brent = 0, nint = 0, skipped = 0;
    // the sides' descriptors stay in global memory (read with uniform
    // addresses): two copies as kernel arguments, picked per fiber, cost
    // registers and spills
    const uint32_t nf0 = sides[0].nf;
    const uint32_t total = nf0 + (nsides > 1 ? sides[1].nf : 0u);
    // Fibers cost very different amounts (a fiber through a sphere against
    // one that misses every triangle), so each wavefront takes the next
    // fiber from a device counter when it is done, instead of a fixed share.
    // The sorted fiber order is cut into one contiguous range per XCD
    // (blocks go to the XCDs round-robin), each with its own counter
    // (queue[0..kXcds)): the wavefronts of an XCD take neighbouring fibers,
    // whose candidate triangles overlap, so they share that XCD's L2; an
    // XCD whose range is done takes fibers from the next ranges.
    const uint32_t xcd = blockIdx.x % kXcds;
    uint32_t tried = 0;  // ranges found empty (uniform)
    for (;;) {
        uint32_t w = 0xffffffffu;
        while (tried < kXcds) {
            const uint32_t r = (xcd + tried) % kXcds;
            const uint32_t r0 = (uint32_t)((uint64_t)total * r / kXcds);
            const uint32_t r1 = (uint32_t)((uint64_t)total * (r + 1) / kXcds);
            uint32_t j = 0;
            if (lane == 0) j = atomicAdd(queue + r, 1u);
            j = uniform_u((uint32_t)__shfl((int)j, 0, 64));
            if (j < r1 - r0) {
                w = r0 + j;
                break;
            }
            ++tried;
        }
        if (w == 0xffffffffu) break;
        const bool second = w >= nf0;
        const PushSide& S = sides[second ? 1 : 0];
        const MeshView pm = S.pm;
        const double* __restrict__ acc = S.acc;
        const int dir = S.dir;
        const PushLists& PL = S.PL;  // read where used (global memory, uniform)
        const uint32_t slot = second ? w - nf0 : w;  // position in the side's fiber order
        const uint32_t src = uniform_u(S.perm[slot]);
        const double* __restrict__ fib = S.fib;
        const FibX f = make_fiber(fib + 6 * (size_t)src);
        const double across = (dir == 0) ? f.p1.y : f.p1.x;
        // search box (batchpushcutter.cpp:170-181), in (across, z)
        const double a0 = across - R, a1 = across + R, v0 = f.p1.z, v1 = f.p1.z + L;
        const float qa0 = __double2float_rd(a0), qa1 = __double2float_ru(a1);
        const float qv0 = __double2float_rd(v0), qv1 = __double2float_ru(v1);
        const bool in_lds = PL.rcap == 0;
        const uint32_t cap = in_lds ? kLdsList : PL.rcap;
        PEntry* EL = in_lds ? lmem[wv] : PL.big + (size_t)slot * PL.rcap;  // where the list is built
        uint32_t n = 0;  // list size (kept by lane 0, broadcast when needed)
        bool ovf = false;
        uint32_t qn = 0;
        auto batch = [&](uint32_t nb) {
            const bool has = lane < nb;
            const uint32_t k = has ? qbuf[lane] : 0u;
            PIval iv;
            piv_init(iv);
            if constexpr (K == K_BULL) {
                CutterP c = cd.c;
                c.kind = K_BULL;
                TriX t;
                if (has) t = load_tri(pm.tri, k);
                // 1. the solves the pair needs (edges that are not horizontal,
                //    can reach the segment and have a sign change)
                uint32_t need = 0, ntask = 0;
                for (int e = 0; e < 3; ++e) {
                    bool nd = false;
                    double lo = 0.0, hi = 0.0;
                    if (has) {
                        const V3 p1 = tri_vtx(t, e), p2 = tri_vtx(t, (e + 1) % 3);
                        PIval scratch;
                        piv_init(scratch);
                        AEll el;
                        nd = !horiz_edge_push(c, f, scratch, p1, p2) && bull_edge_prep(c, f, p1, p2, el) &&
                             bull_edge_bracket(el, f, lo, hi);
                    }
                    bmem[wv][2 * e][lane] = lo;
                    bmem[wv][2 * e + 1][lane] = hi;
                    need |= nd ? (1u << e) : 0u;
                    const unsigned long long bal = __ballot(nd);
                    const uint32_t pos = ntask + 2u * (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                    if (nd) {
                        tmem[wv][pos] = (lane << 3) | ((uint32_t)e << 1);
                        tmem[wv][pos + 1] = (lane << 3) | ((uint32_t)e << 1) | 1u;
                    }
                    ntask += 2u * (uint32_t)__popcll(bal);
                }
                wave_lds_sync();
                // 2. the listed solves: every lane keeps one going and takes
                //    the next listed solve when its own ends, one error
                //    evaluation per step (brent_step)
                {
                    uint32_t cursor = 0;  // next solve to hand out
                    bool busy = false;
                    uint32_t td = 0;
                    BrentState B;
                    AEll el;
                    for (;;) {
                        const bool want = !busy;
                        const unsigned long long wb = __ballot(want);
                        const uint32_t tk = cursor + (uint32_t)__popcll(wb & ((1ull << lane) - 1ull));
                        if (want && tk < ntask) {
                            td = tmem[wv][tk];
                            const uint32_t pl = td >> 3, e = (td >> 1) & 3u;
                            const TriX tt = load_tri(pm.tri, qbuf[pl]);
                            bull_edge_prep(c, f, tri_vtx(tt, (int)e), tri_vtx(tt, ((int)e + 1) % 3), el);
                            const double lo = bmem[wv][2 * e][pl], hi = bmem[wv][2 * e + 1][pl];
                            if (td & 1u) brent_start(B, hi - 4.0, lo);
                            else brent_start(B, lo, hi);
                            busy = true;
                        }
                        cursor += (uint32_t)__popcll(wb);
                        if (!__any(busy)) break;
                        if (busy && brent_step(B, aell_error(el, brent_x(B)), 3E-16, 1e-10)) {
                            const uint32_t pl = td >> 3, e = (td >> 1) & 3u;
                            rmem[wv][2 * e + (td & 1u)][pl] = B.b;
                            brent += B.it;
                            busy = false;
                        }
                    }
                }
                wave_lds_sync();
                // 3. the pair in the reference's order with the roots
                //    (millingcutter.cpp:377-382, 277-301)
                if (has) {
                    base_vertex_push(c, f, iv, t);
                    simple_facet_push(c, f, iv, t);
                    for (int e = 0; e < 3; ++e) {
                        const V3 p1 = tri_vtx(t, e), p2 = tri_vtx(t, (e + 1) % 3);
                        if (horiz_edge_push(c, f, iv, p1, p2)) continue;
                        shaft_edge_push(c, f, iv, p1, p2);
                        if (need & (1u << e)) {
                            AEll el;
                            bull_edge_prep(c, f, p1, p2, el);
                            bull_edge_finish(c, f, iv, p1, p2, el, rmem[wv][2 * e][lane], rmem[wv][2 * e + 1][lane]);
                        }
                    }
                }
                wave_lds_sync();
            } else {
                if (has) {
                    const TriX t = load_tri(pm.tri, k);
                    int it = 0;
                    push_pair<K>(cd, f, t, iv, it);
                    brent += it;
                }
            }
            const bool keep = has && !piv_empty(iv);
            if (keep) ebuf[lane] = pentry_of(iv);
            wave_lds_sync();
            // The batch's intervals merge first among themselves: sorted by
            // (lower end, rank), a connected component starts where the lower
            // end exceeds every upper end before it (touching ends merge, as
            // Interval::outside is strict).  A component's lower end and CC
            // are its first member's (the lowest, earliest on ties), its upper
            // end and CC those of its highest member, earliest on ties: what
            // inserting the members one by one in rank order would keep, since
            // an existing entry keeps its extremes on ties.  Lane 0 then
            // inserts the components (disjoint, so their order is free).
            double key = keep ? iv.lo : INFINITY;
            uint32_t idx = lane;
            for (uint32_t kk = 2; kk <= 64; kk <<= 1) {
                for (uint32_t jj = kk >> 1; jj > 0; jj >>= 1) {
                    const double ok = __shfl_xor(key, (int)jj, 64);
                    const uint32_t oi = (uint32_t)__shfl_xor((int)idx, (int)jj, 64);
                    const bool up = (lane & kk) == 0, low = (lane & jj) == 0;
                    const bool less = (ok < key) || (ok == key && oi < idx);
                    const bool take = (up == low) ? less : !less;
                    key = take ? ok : key;
                    idx = take ? oi : idx;
                }
            }
            const bool valid = key < INFINITY;
            const double hi_s = __shfl(keep ? iv.hi : -INFINITY, (int)idx, 64);
            // upper ends before each position: an exclusive prefix max
            double pmax = valid ? hi_s : -INFINITY;
            for (int off = 1; off < 64; off <<= 1) {
                const double o = __shfl_up(pmax, off, 64);
                if ((int)lane >= off) pmax = fmax(pmax, o);
            }
            double before = __shfl_up(pmax, 1, 64);
            if (lane == 0) before = -INFINITY;
            const bool start = valid && key > before;
            const unsigned long long sb = __ballot(start);
            const uint32_t comp = (uint32_t)__popcll(sb & ((2ull << lane) - 1ull)) - 1u;
            const uint32_t ncomp = (uint32_t)__popcll(sb);
            unsigned long long rest = sb;
            for (uint32_t c = 0; c < ncomp; ++c) {
                const uint32_t first = (uint32_t)__ffsll((long long)rest) - 1u;
                rest &= rest - 1;
                const bool in_c = valid && comp == c;
                const double mx = wave_max(in_c ? hi_s : -INFINITY);
                const uint32_t top = (uint32_t)wave_min_t<double>(in_c && hi_s == mx ? (double)idx : 64.0);
                const uint32_t lead = uniform_u((uint32_t)__shfl((int)idx, (int)first, 64));
                if (lane == 0 && !ovf) {
                    PEntry m = ebuf[lead];
                    const PEntry& u = ebuf[top];
                    m.hi = u.hi;
                    m.ucc[0] = u.ucc[0]; m.ucc[1] = u.ucc[1]; m.ucc[2] = u.ucc[2];
                    m.uty = u.uty;
                    if (!plist_insert(EL, n, cap, m)) ovf = true;
                }
            }
            wave_lds_sync();
            pairs += has ? 1 : 0;
        };
        // the walk pauses whenever 64 candidates are queued, so the batch
        // code is emitted once
        // with skip, the walk also needs the box's reach box to meet the
        // fiber's reach query, a superset of push_bound's own tests for
        // every triangle of the box: across +- (R + margin), highest vertex
        // at or above the fiber less the slack
        const bool prune = skip && K != K_COMPOSITE;
        const double ma = R + 2e-3 + 1e-8 * (fabs(across) + R + 1.0);
        const float ra0 = __double2float_rd(across - ma), ra1 = __double2float_ru(across + ma);
        const float rz0 = __double2float_rd(f.p1.z - 2e-4 - 3e-9 * (fabs(f.p1.z) + 1.0));
        const float4* __restrict__ reach = S.reach;
        auto hit = [&](uint32_t i) -> bool {
            return box_hits(pm.tree.box[i], qa0, qv0, qa1, qv1) &&
                   (!prune || box_hits(reach[i], ra0, rz0, ra1, INFINITY));
        };
        BoxWalk walk;
        for (;;) {
            walk_boxes_resumable_if(pm.tree, hit, walk, [](uint32_t) {}, [&](uint32_t b, uint32_t) -> bool {
                const uint32_t k = b * kLeaf + lane;
                bool cand = false;
                if (lane < (uint32_t)kLeaf && k < pm.ntri) {
                    if (superset) {
                        const double* q = acc + 4 * (size_t)k;  // needs max >= A and min <= B per axis
                        cand = a1 >= q[0] && a0 <= q[1] && v1 >= q[2] && v0 <= q[3];
                    } else {
                        const TriX t = load_tri(pm.tri, k);
                        const double ta0 = fmin(fmin(t.p[0].x, t.p[1].x), t.p[2].x);
                        const double ta1 = fmax(fmax(t.p[0].x, t.p[1].x), t.p[2].x);
                        const double tb0 = fmin(fmin(t.p[0].y, t.p[1].y), t.p[2].y);
                        const double tb1 = fmax(fmax(t.p[0].y, t.p[1].y), t.p[2].y);
                        const double lo_a = (dir == 0) ? tb0 : ta0, hi_a = (dir == 0) ? tb1 : ta1;
                        const double lo_z = fmin(fmin(t.p[0].z, t.p[1].z), t.p[2].z);
                        cand = !(hi_a < a0 || lo_a > a1 || t.maxz < v0 || lo_z > v1);
                    }
                }
                if (skip) {
                    // the list's size is lane 0's; its entries are in LDS,
                    // written before the last batch's closing wave_lds_sync
                    // (the retry pass, whose list is in global memory, tests
                    // only whether a pair can reach its fiber)
                    const uint32_t nl = in_lds ? uniform_u((uint32_t)__shfl((int)n, 0, 64)) : 0u;
                    if (cand) {
                        const double* r = pm.tri + (size_t)kRec * k;
                        const double rb[3] = {r[13], r[14], r[15]};
                        const int ca = dir == 0 ? 1 : 0;  // coordinate across the fiber
                        const double c0 = fmin(fmin(r[ca], r[3 + ca]), r[6 + ca]);
                        const double c1 = fmax(fmax(r[ca], r[3 + ca]), r[6 + ca]);
                        double tl = 0.0, tu = 0.0;
                        const int bd = push_bound(cd.c, rb, r[12], c0, c1, f.p1.z, across, dir == 0 ? f.p1.x : f.p1.y,
                                                  dir == 0 ? f.p2.x : f.p2.y, tl, tu);
                        bool drop = bd == 1;
                        if (bd == 2) {
                            // the last entry starting below tl (the list is
                            // sorted by lower end and disjoint)
                            uint32_t lo = 0, hi = nl;
                            while (lo < hi) {
                                const uint32_t mid = (lo + hi) >> 1;
                                if (EL[mid].lo < tl) lo = mid + 1;
                                else hi = mid;
                            }
                            drop = lo > 0 && tu < EL[lo - 1].hi;
                        }
                        skipped += drop ? 1 : 0;
                        cand = !drop;
                    }
                }
                const unsigned long long bal = __ballot(cand);
                const uint32_t pos = qn + __popcll(bal & ((1ull << lane) - 1ull));
                if (cand) qbuf[pos] = k;
                qn += __popcll(bal);
                wave_lds_sync();
                return qn < 64;
            });
            if (qn >= 64 || (walk.done && qn > 0)) {
                const uint32_t nb = qn < 64 ? qn : 64u;
                batch(nb);
                // a list that outgrew its capacity is redone in a retry pass
                if (__shfl((int)ovf, 0, 64) != 0) break;
                const uint32_t rest = qn - nb;
                const uint32_t moved = (lane < rest) ? qbuf[64 + lane] : 0u;
                wave_lds_sync();
                if (lane < rest) qbuf[lane] = moved;
                wave_lds_sync();
                qn = rest;
            }
            if (walk.done && qn == 0) break;
        }
        const bool of = __shfl((int)ovf, 0, 64) != 0;
        const uint32_t nn = uniform_u(__shfl(n, 0, 64));
        if (in_lds) {
            // copy the list into the pool (n and the offset are lane 0's)
            unsigned long long o = 0;
            if (lane == 0 && !of) o = atomicAdd(PL.used, (unsigned long long)nn);
            o = __shfl(o, 0, 64);
            const bool fits = !of && o + nn <= PL.pool_cap;
            wave_lds_sync();
            if (fits)
                for (uint32_t j = lane; j < nn; j += 64) PL.pool[o + j] = lmem[wv][j];
            wave_lds_sync();
            if (lane == 0) {
                PL.n[src] = fits ? nn : 0u;
                PL.off[src] = fits ? (uint64_t)o : 0ull;
                if (!fits) PL.retry[atomicAdd(PL.retry_n, 1u)] = src;
                nint += fits ? nn : 0u;
            }
        } else if (lane == 0) {
            PL.n[src] = of ? 0u : nn;
            PL.off[src] = kBigList | ((uint64_t)slot * PL.rcap);
            if (of) atomicAdd(PL.retry_n, 1u);
            nint += of ? 0u : nn;
        }
    }
    if (cnt) {
        wave_add(&cnt->pairs, pairs);
        wave_add(&cnt->intervals, nint);
        wave_add(&cnt->brent, brent);
        wave_add(&cnt->skipped, skipped);
    }
}

// Output gather: fiber f's entries to rows offs[f]..offs[f+1]-1 of
// lohi (2 per row) and cc (lower xyz, type, upper xyz, type).
__global__ void k_push_gather(PushLists PL, const uint64_t* __restrict__ offs, uint32_t nf, double* __restrict__ lohi,
                              double* __restrict__ cc) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const uint64_t at = PL.off[f];
    const PEntry* E = (at & kBigList) ? PL.big + (at & ~kBigList) : PL.pool + at;
    const uint64_t o = offs[f];
    const uint32_t n = PL.n[f];
    for (uint32_t j = 0; j < n; ++j) {
        const PEntry& p = E[j];
        lohi[2 * (o + j)] = p.lo;
        lohi[2 * (o + j) + 1] = p.hi;
        double* c = cc + 8 * (o + j);
        c[0] = p.lcc[0]; c[1] = p.lcc[1]; c[2] = p.lcc[2]; c[3] = (double)p.lty;
        c[4] = p.ucc[0]; c[5] = p.ucc[1]; c[6] = p.ucc[2]; c[7] = (double)p.uty;
    }
}

}  // namespace ocl_amd
