// Device-side cutter description.
//
// Mirrors the MillingCutter family (src/cutters/*.cpp) as plain data: the
// constructor-derived constants (radius, length, xy_normal_length,
// normal_length, center_height, corner radii, cone angle and its tangent) are
// filled on the host, in FP64, by CutterP::make_* so that every consumer
// (exact FP64 path, fast FP32/FP64 path, push path) sees the same values.
#pragma once

#include <cmath>
#include <cstdint>

namespace ocl_amd {

enum CutterKind : int { K_CYL = 0, K_BALL = 1, K_BULL = 2, K_CONE = 3, K_COMPOSITE = 4 };

// CCType numbering of src/geo/ccpoint.hpp:33-39
enum CCType : int {
    CC_NONE = 0, CC_VERTEX, CC_VERTEX_CYL, CC_EDGE, CC_EDGE_HORIZ, CC_EDGE_SHAFT,
    CC_EDGE_HORIZ_CYL, CC_EDGE_HORIZ_TOR, CC_EDGE_BALL, CC_EDGE_POS, CC_EDGE_NEG,
    CC_EDGE_CYL, CC_EDGE_CONE, CC_EDGE_CONE_BASE, CC_FACET, CC_FACET_TIP, CC_FACET_CYL, CC_ERROR
};

// Contact feature codes carried from the selection pass to the FP64 refine:
// low nibble = feature, bits 4..5 = composite sub-cutter.
enum Feature : int { FT_FACET = 0, FT_FACET_RIM = 1, FT_VTX0 = 2, FT_EDGE0 = 5, FT_NONE = 15 };

struct CutterP {
    int kind;
    double diameter, radius, length;
    double xyl, nl, ch;  // xy_normal_length, normal_length, center_height
    double r1, r2;       // BullCutter radius1 (shaft) and radius2 (corner)
    double angle, tan_a; // ConeCutter half-angle and tan(angle)

    // cylcutter.cpp:39-46
    static CutterP cyl(double d, double l) {
        CutterP c{};
        c.kind = K_CYL; c.diameter = d; c.radius = d / 2.0; c.length = l;
        c.xyl = c.radius; c.nl = 0.0; c.ch = 0.0;
        return c;
    }
    // ballcutter.cpp:38-45
    static CutterP ball(double d, double l) {
        CutterP c{};
        c.kind = K_BALL; c.diameter = d; c.radius = d / 2.0; c.length = l;
        c.nl = c.radius; c.xyl = 0.0; c.ch = c.radius;
        return c;
    }
    // bullcutter.cpp:38-47
    static CutterP bull(double d, double r, double l) {
        CutterP c{};
        c.kind = K_BULL; c.diameter = d; c.radius = d / 2.0;
        c.r1 = d / 2.0 - r; c.r2 = r; c.length = l;
        c.xyl = c.r1; c.nl = c.r2; c.ch = c.r2;
        return c;
    }
    // conecutter.cpp:39-49
    static CutterP cone(double d, double a, double l) {
        CutterP c{};
        c.kind = K_CONE; c.diameter = d; c.radius = d / 2.0; c.angle = a;
        c.tan_a = std::tan(a);
        c.length = c.radius / c.tan_a + l;
        c.ch = c.radius / c.tan_a;
        c.xyl = c.radius; c.nl = 0.0;
        return c;
    }
};

constexpr int kMaxSub = 4;

struct CutterDev {
    CutterP c;        // the cutter itself (for a composite: radius/length only)
    int nsub;         // composite: number of sub-cutters
    CutterP sub[kMaxSub];
    double rlim[kMaxSub], hlim[kMaxSub], zoff[kMaxSub];
};

}  // namespace ocl_amd
