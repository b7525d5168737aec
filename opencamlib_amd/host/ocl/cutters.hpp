// MillingCutter family of the drop-in host API.
//
// Constructors and accessors follow the reference (src/cutters/*.hpp):
// CylCutter(d, L) (cylcutter.cpp:39-46), BallCutter(d, L)
// (ballcutter.cpp:38-45), BullCutter(d, r, L) (bullcutter.cpp:38-47),
// ConeCutter(d, angle, l = 10) with length = r / tan(angle) + l
// (conecutter.cpp:39-49), CompositeCutter::addCutter(cutter, radius, height,
// zoffset) (compositecutter.cpp:43-53) and the composite presets
// (compositecutter.cpp:296-392).  The cutters carry no geometry code: each
// one describes itself to the device as an ocl_cutter_desc, and the HIP
// kernels evaluate the contacts.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ocl_hip.h"

namespace ocl {

class MillingCutter {
  public:
    virtual ~MillingCutter() = default;
    double getDiameter() const { return diameter; }
    double getRadius() const { return radius; }
    double getLength() const { return length; }
    virtual std::string str() const = 0;
    // descriptor handed to the device kernels
    virtual ocl_cutter_desc desc() const = 0;

  protected:
    double diameter = 0, radius = 0, length = 0;
};

class CylCutter : public MillingCutter {
  public:
    CylCutter(double d, double l);
    std::string str() const override;
    ocl_cutter_desc desc() const override;
};

class BallCutter : public MillingCutter {
  public:
    BallCutter(double d, double l);
    std::string str() const override;
    ocl_cutter_desc desc() const override;
};

class BullCutter : public MillingCutter {
  public:
    BullCutter(double d, double r, double l);
    std::string str() const override;
    ocl_cutter_desc desc() const override;
    double getCornerRadius() const { return corner; }

  private:
    double corner;
};

class ConeCutter : public MillingCutter {
  public:
    ConeCutter(double d, double angle, double l = 10);
    std::string str() const override;
    ocl_cutter_desc desc() const override;
    double getAngle() const { return angle; }

  private:
    double angle, extra;
};

// Piecewise cutter: sub-cutter n is valid out to radius radiusvec[n] and up to
// height heightvec[n], shifted down by zoffset[n].  The device takes up to
// four sub-cutters, none of them composite.
class CompositeCutter : public MillingCutter {
  public:
    CompositeCutter() = default;
    // the sub-cutter is copied (the reference keeps a pointer)
    void addCutter(const MillingCutter& c, double r, double h, double zoff);
    void setLength(double l) { length = l; }
    std::string str() const override;
    ocl_cutter_desc desc() const override;

  protected:
    std::vector<ocl_cutter_desc> subs;
    std::vector<double> radiusvec, heightvec, zoffset;
};

class CompCylCutter : public CompositeCutter {
  public:
    CompCylCutter(double d, double l);
};
class CompBallCutter : public CompositeCutter {
  public:
    CompBallCutter(double d, double l);
};
class CylConeCutter : public CompositeCutter {
  public:
    CylConeCutter(double diam1, double diam2, double angle);
};
class BallConeCutter : public CompositeCutter {
  public:
    BallConeCutter(double diam1, double diam2, double angle);
};
class BullConeCutter : public CompositeCutter {
  public:
    BullConeCutter(double diam1, double radius1, double diam2, double angle);
};
class ConeConeCutter : public CompositeCutter {
  public:
    ConeConeCutter(double diam1, double angle1, double diam2, double angle2);
};

}  // namespace ocl
