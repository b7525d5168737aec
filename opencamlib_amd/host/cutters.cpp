// MillingCutter family of the drop-in host API (see ocl/cutters.hpp).
#include <cmath>
#include <cstring>
#include <sstream>

#include "ocl/cutters.hpp"
#include "ocl/geometry.hpp"

namespace ocl {

static ocl_cutter_desc simple_desc(int type, double d, double l, double r, double angle) {
    ocl_cutter_desc s;
    std::memset(&s, 0, sizeof(s));
    s.c.type = type;
    s.c.diameter = d;
    s.c.length = l;
    s.c.corner_radius = r;
    s.c.angle = angle;
    return s;
}

static void need(bool ok, const char* what) {
    if (!ok) throw Error(what);
}

CylCutter::CylCutter(double d, double l) {
    need(d > 0 && l > 0, "CylCutter needs diameter > 0 and length > 0");
    diameter = d;
    radius = d / 2;
    length = l;
}
std::string CylCutter::str() const {
    std::ostringstream o;
    o << "CylCutter (d=" << diameter << ", L=" << length << ")";
    return o.str();
}
ocl_cutter_desc CylCutter::desc() const { return simple_desc(OCL_CUTTER_CYL, diameter, length, 0, 0); }

BallCutter::BallCutter(double d, double l) {
    need(d > 0 && l > 0, "BallCutter needs diameter > 0 and length > 0");
    diameter = d;
    radius = d / 2;
    length = l;
}
std::string BallCutter::str() const {
    std::ostringstream o;
    o << "BallCutter (d=" << diameter << ", r=" << radius << ", L=" << length << ")";
    return o.str();
}
ocl_cutter_desc BallCutter::desc() const { return simple_desc(OCL_CUTTER_BALL, diameter, length, 0, 0); }

BullCutter::BullCutter(double d, double r, double l) : corner(r) {
    need(d > 0 && l > 0, "BullCutter needs diameter > 0 and length > 0");
    need(r > 0 && d / 2 - r > 0, "BullCutter needs 0 < corner radius < diameter/2");
    diameter = d;
    radius = d / 2;
    length = l;
}
std::string BullCutter::str() const {
    std::ostringstream o;
    o << "BullCutter (d=" << diameter << ", r1=" << radius - corner << ", r2=" << corner << ", L=" << length << ")";
    return o.str();
}
ocl_cutter_desc BullCutter::desc() const { return simple_desc(OCL_CUTTER_BULL, diameter, length, corner, 0); }

ConeCutter::ConeCutter(double d, double a, double l) : angle(a), extra(l) {
    need(d > 0, "ConeCutter needs diameter > 0");
    need(a > 0 && a < 1.5707963267948966, "ConeCutter angle must be in (0, pi/2)");
    diameter = d;
    radius = d / 2;
    length = radius / std::tan(angle) + l;  // conecutter.cpp:44
}
std::string ConeCutter::str() const {
    std::ostringstream o;
    o << "ConeCutter (d=" << diameter << ", angle=" << angle << ", L=" << length << ")";
    return o.str();
}
ocl_cutter_desc ConeCutter::desc() const { return simple_desc(OCL_CUTTER_CONE, diameter, extra, 0, angle); }

void CompositeCutter::addCutter(const MillingCutter& c, double r, double h, double zoff) {
    const ocl_cutter_desc d = c.desc();
    need(d.c.type != OCL_CUTTER_COMPOSITE, "CompositeCutter: a sub-cutter cannot be composite");
    need(subs.size() < 4, "CompositeCutter: at most 4 sub-cutters");
    subs.push_back(d);
    radiusvec.push_back(r);
    heightvec.push_back(h);
    zoffset.push_back(zoff);
    if (r > radius) {  // compositecutter.cpp:48-51
        radius = r;
        diameter = 2 * r;
    }
}

std::string CompositeCutter::str() const {
    std::ostringstream o;
    o << "CompositeCutter with " << subs.size() << " cutters";
    for (size_t i = 0; i < subs.size(); ++i)
        o << "\n " << i << ": type " << subs[i].c.type << " d=" << subs[i].c.diameter << " radius=" << radiusvec[i]
          << " height=" << heightvec[i] << " zoffset=" << zoffset[i];
    return o.str();
}

ocl_cutter_desc CompositeCutter::desc() const {
    need(!subs.empty(), "CompositeCutter has no sub-cutters");
    ocl_cutter_desc d;
    std::memset(&d, 0, sizeof(d));
    d.c.type = OCL_CUTTER_COMPOSITE;
    d.c.diameter = diameter;
    d.c.length = length;
    d.nsub = (int32_t)subs.size();
    for (size_t i = 0; i < subs.size(); ++i) {
        d.sub[i] = subs[i].c;
        d.radius_limit[i] = radiusvec[i];
        d.height_limit[i] = heightvec[i];
        d.zoffset[i] = zoffset[i];
    }
    return d;
}

// ------------------------------------------------ composite presets
// Same sub-cutters, validity radii, heights and offsets as the reference
// presets (compositecutter.cpp:296-392); the fixed shaft lengths (20, 30)
// are the reference's.

CompCylCutter::CompCylCutter(double d, double l) {
    addCutter(CylCutter(d, l), d / 2, l, 0.0);
    length = l;
}

CompBallCutter::CompBallCutter(double d, double l) {
    addCutter(BallCutter(d, l), d / 2, l, 0.0);
    length = l;
}

CylConeCutter::CylConeCutter(double diam1, double diam2, double angle) {
    const double t = std::tan(angle);
    const double cone_off = -(diam1 / 2) / t;
    const double cyl_h = 0.0;
    const double cone_h = (diam2 / 2) / t + cone_off;
    addCutter(CylCutter(diam1, 1), diam1 / 2, cyl_h, 0.0);
    addCutter(ConeCutter(diam2, angle), diam2 / 2, cone_h, cone_off);
    addCutter(CylCutter(diam2, 20), diam2 / 2, (diam2 / 2) / t + 20, cone_h);
    length = cyl_h + cone_h + 10;
}

BallConeCutter::BallConeCutter(double diam1, double diam2, double angle) {
    const double r1 = diam1 / 2, r2 = diam2 / 2;
    const double rc = r1 * std::cos(angle);  // where the cone meets the ball tangentially
    const double h1 = r1 - std::sqrt(r1 * r1 - rc * rc);
    const double cone_off = -(rc / std::tan(angle) - h1);
    const double h2 = r2 / std::tan(angle) + cone_off;
    addCutter(BallCutter(diam1, 1), rc, h1, 0.0);
    addCutter(ConeCutter(diam2, angle), r2, h2, cone_off);
    addCutter(CylCutter(diam2, 20), r2, h2 + 20, h2);
    length = 30;
}

BullConeCutter::BullConeCutter(double diam1, double radius1, double diam2, double angle) {
    const double drop = radius1 * std::sin(angle);  // contact below the torus ring
    const double ring = std::sqrt(radius1 * radius1 - drop * drop);
    const double rc = diam1 / 2 - radius1 + ring;
    const double cone_off = -(rc / std::tan(angle) - (radius1 - drop));
    const double h1 = radius1 - drop;
    const double h2 = (diam2 / 2) / std::tan(angle) + cone_off;
    addCutter(BullCutter(diam1, radius1, 1.0), rc, h1, 0.0);
    addCutter(ConeCutter(diam2, angle), diam2 / 2, h2, cone_off);
    addCutter(CylCutter(diam2, 20), diam2 / 2, h2 + 20, h2);
    length = 30;
}

ConeConeCutter::ConeConeCutter(double diam1, double angle1, double diam2, double angle2) {
    const double h1 = (diam1 / 2) / std::tan(angle1);
    const double cone_off = -((diam1 / 2) / std::tan(angle2) - h1);
    const double h2 = (diam2 / 2) / std::tan(angle2) + cone_off;
    addCutter(ConeCutter(diam1, angle1), diam1 / 2, h1, 0.0);
    addCutter(ConeCutter(diam2, angle2), diam2 / 2, h2, cone_off);
    addCutter(CylCutter(diam2, 20), diam2 / 2, h2 + 20, h2);
    length = 30;
}

}  // namespace ocl
