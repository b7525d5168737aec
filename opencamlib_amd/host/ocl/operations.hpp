// Operation classes of the drop-in host API: BatchDropCutter, PointDropCutter,
// BatchPushCutter, PathDropCutter, AdaptivePathDropCutter, with the Fiber /
// Interval value types.
//
// The public methods are the reference's (src/algo/operation.hpp:43-138,
// src/dropcutter/batchdropcutter.hpp:48-79, pointdropcutter.hpp,
// pathdropcutter.hpp, adaptivepathdropcutter.hpp,
// src/algo/batchpushcutter.hpp:46-84, fiber.hpp:35-86, interval.hpp:37-94).
// The work is done on the GPU through the C ABI of include/ocl_hip.h; the
// classes hold a device mesh handle built in setSTL.
//
// Behaviour that differs from the reference, on purpose:
//   - setThreads() only records the value: the device decides its own
//     parallelism.
//   - getCalls() is the number of (point, triangle) pairs the device
//     evaluated fully, not the reference's dropCutter() call count; the
//     device culls pairs that cannot change the result.
//   - Precision: setPrecision(OCL_PREC_*) selects the device mode; the
//     default is OCL_PREC_MIXED (FP32 selection, FP64 contact of the winner,
//     exact contact types).  OCL_PREC_EXACT64 reproduces the reference's FP64
//     fold on every pair.
//   - Push: each fiber's intervals come back sorted by lower end (the
//     reference lists them in merge order).  Intervals a fiber held before
//     run() are merged with Fiber::addInterval semantics.
//   - Failures throw ocl::Error with the device library's message.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ocl/cutters.hpp"
#include "ocl/geometry.hpp"
#include "ocl_hip.h"

namespace ocl {

struct Interval {
    double upper = 0, lower = 0;
    CCPoint upper_cc, lower_cc;
    bool in_weave = false;
    Interval() = default;
    Interval(double l, double u);
    // the first update sets both ends (interval.cpp:61-93)
    void updateUpper(double t, const CCPoint& p);
    void updateLower(double t, const CCPoint& p);
    void update(double t, const CCPoint& p) {
        updateUpper(t, p);
        updateLower(t, p);
    }
    bool outside(const Interval& i) const { return lower > i.upper || upper < i.lower; }
    bool inside(const Interval& i) const { return lower > i.lower && upper < i.upper; }
    bool empty() const { return lower == 0.0 && upper == 0.0; }
    std::string str() const;
};

struct Fiber {
    Point p1, p2, dir;
    std::vector<Interval> ints;
    Fiber() = default;
    // p1, p2 at one height (fiber.cpp:29-39)
    Fiber(const Point& a, const Point& b);
    void addInterval(const Interval& i);
    bool contains(const Interval& i) const;
    bool missing(const Interval& i) const;
    double tval(const Point& p) const;
    Point point(double t) const { return p1 + (p2 - p1) * t; }
    bool empty() const { return ints.empty(); }
    unsigned int size() const { return (unsigned int)ints.size(); }
    Point upperCLPoint(unsigned int n) const { return point(ints.at(n).upper); }
    Point lowerCLPoint(unsigned int n) const { return point(ints.at(n).lower); }
    bool operator==(const Fiber& o) const { return p1 == o.p1 && p2 == o.p2; }
    std::string str() const;
};

// Owns a device mesh (ocl_mesh) built from an STLSurf.
class DeviceSurface {
  public:
    DeviceSurface() = default;
    ~DeviceSurface();
    DeviceSurface(const DeviceSurface&) = delete;
    DeviceSurface& operator=(const DeviceSurface&) = delete;
    void build(const STLSurf& s, unsigned bucket);
    ocl_mesh* handle() const { return mesh; }
    // triangle index (input order) -> triangle, for the debug queries
    const std::vector<Triangle>& triangles() const { return tris; }
    // Triangles the reference's KD search returns for a query box (a0..a1 on
    // the first search axis, b0..b1 on the second; dimset 0 XY, 1 YZ, 2 XZ),
    // in its order.  The KD order and acceptance bounds of a dimension set
    // are built on the first query and kept until the next build.
    std::vector<Triangle> kd_query(int dimset, unsigned bucket, double a0, double a1, double b0, double b1);

  private:
    struct KdOrder {
        bool built = false;
        unsigned bucket = 0;
        std::vector<uint32_t> order;  // rank -> triangle index
        std::vector<double> acc;      // per rank: the search's acceptance bounds
    };
    ocl_mesh* mesh = nullptr;
    std::vector<Triangle> tris;
    KdOrder kd[3];
};

class Operation {
  public:
    Operation() = default;
    virtual ~Operation() = default;
    Operation(const Operation&) = delete;
    Operation& operator=(const Operation&) = delete;

    // copies the surface to the device with the current bucket size (the
    // reference builds its KD-tree here, so setBucketSize must come first)
    virtual void setSTL(const STLSurf& s);
    // non-owning, as in the reference: the cutter must outlive run()
    virtual void setCutter(const MillingCutter* c) { cutter = c; }
    void setThreads(unsigned int n) { nthreads = n; }
    unsigned int getThreads() const { return nthreads; }
    void setBucketSize(unsigned int s) { bucketSize = s; }
    unsigned int getBucketSize() const { return bucketSize; }
    int getCalls() const { return (int)nCalls; }
    void setPrecision(int p) { precision = p; }
    int getPrecision() const { return precision; }
    // counters of the last device call
    const ocl_stats& getStats() const { return stats; }
    virtual void run() = 0;

  protected:
    void require_ready() const;
    const STLSurf* surf = nullptr;
    const MillingCutter* cutter = nullptr;
    unsigned int nthreads = 1, bucketSize = 1;
    unsigned long long nCalls = 0;
    int precision = OCL_PREC_MIXED;
    ocl_stats stats{};
    DeviceSurface dev;
};

class BatchDropCutter : public Operation {
  public:
    void appendPoint(const CLPoint& p) { clpoints.push_back(p); }
    void run() override;
    std::vector<CLPoint> getCLPoints() const { return clpoints; }
    void clearCLPoints() { clpoints.clear(); }
    // triangles the reference's XY KD search returns for a cutter at cl, in
    // its order (batchdropcutter_py.hpp:46-55)
    std::vector<Triangle> getTrianglesUnderCutter(const CLPoint& cl, const MillingCutter& c);

  protected:
    std::vector<CLPoint> clpoints;
};

// One point at a time (pointdropcutter.cpp:56-81): the single-point latency
// path, one device launch per call.
class PointDropCutter : public Operation {
  public:
    void run() override {}
    void run(CLPoint& cl);
};

class BatchPushCutter : public Operation {
  public:
    // the direction can be chosen before or after setSTL (the device keeps
    // one tree per direction)
    void setXDirection() { dir = 0; }
    void setYDirection() { dir = 1; }
    void appendFiber(const Fiber& f) { fibers.push_back(f); }
    void run() override;
    std::vector<Fiber>* getFibers() { return &fibers; }
    void reset() { fibers.clear(); }
    // lower and upper CL points with their CC points of every non-empty
    // interval (batchpushcutter_py.hpp:37-57)
    std::vector<CLPoint> getCLPoints() const;
    // triangles the reference's KD search returns for fiber f
    // (batchpushcutter_py.hpp:59-84)
    std::vector<Triangle> getOverlapTriangles(const Fiber& f);
    // push only triangles whose boxes overlap the search box (the reference
    // pushes its whole KD-search superset)
    void setExactFilter(bool on) { exact_filter = on; }

  protected:
    int dir = -1;
    bool exact_filter = false;
    std::vector<Fiber> fibers;
};

// Samples each span of a path at the given step and drops the cutter on all
// samples in one device batch (pathdropcutter.cpp:54-80).
class PathDropCutter : public Operation {
  public:
    void setPath(const Path* p) { path = p; }
    void setSampling(double s) { sampling = s; }
    double getSampling() const { return sampling; }
    void setZ(double z) { minimumZ = z; }
    double getZ() const { return minimumZ; }
    void run() override;
    std::vector<CLPoint> getCLPoints() const { return clpoints; }

  protected:
    const Path* path = nullptr;
    double sampling = 0.1, minimumZ = 0.0;
    std::vector<CLPoint> clpoints;
};

// Adaptive sampling (adaptivepathdropcutter.cpp:53-94): a segment is split
// at its middle while its XY length exceeds the sampling step, or while it
// is not flat and longer than the minimum step.  The recursion is evaluated
// level by level, all segments of a level in one device batch; the output is
// the reference's depth-first order.
//
// setChordTolerance(tol > 0) replaces the flatness test by slope-driven
// refinement: every sample carries the slope of z along the path (from the
// device backward pass).  An interval passes when the piecewise cubic
// Hermite curve through its two ends and its middle (heights and slopes)
// stays within tol/2 of the chord (the other half is the margin for what
// the samples cannot see); it is then checked at its quarter points, and
// closes when both lie within tol/2 of the chord.  Otherwise it is split
// at the middle (after a failed check, the quarter points are the halves'
// middles).  Intervals at the minimum step close.  The first samples are
// spaced at most `sampling` apart; each level of intervals is one device
// batch.
class AdaptivePathDropCutter : public PathDropCutter {
  public:
    void setMinSampling(double s) { min_sampling = s; }
    void setCosLimit(double c) { cosLimit = c; }
    void setChordTolerance(double tol) { chord_tol = tol; }
    double getChordTolerance() const { return chord_tol; }
    // recursion levels of the last run (one device batch each)
    int getLevels() const { return levels; }
    void run() override;

  protected:
    void flatness_run();
    void slope_run();
    double min_sampling = 0.01, cosLimit = 0.999, chord_tol = 0.0;
    int levels = 0;
};

// Waterline fibers (waterline.cpp:67-87, 143-178): X fibers at every y and
// Y fibers at every x of a grid spanning the surface's bounding box grown by
// twice the cutter radius, at height z, `sampling` apart, pushed on the
// device, both directions in one launch (ocl_hip_push_xy).  The weave that turns the
// fibers into loops is graph work the reference does with Boost Graph and is
// not built here: getLoops() throws.
class Waterline : public Operation {
  public:
    void setZ(double z) { zh = z; }
    double getZ() const { return zh; }
    void setSampling(double s) { sampling = s; }
    double getSampling() const { return sampling; }
    void setExactFilter(bool on) { exact_filter = on; }
    void run() override;
    void reset() {
        xfibers.clear();
        yfibers.clear();
    }
    const std::vector<Fiber>& getXFibers() const { return xfibers; }
    const std::vector<Fiber>& getYFibers() const { return yfibers; }
    std::vector<std::vector<Point>> getLoops() const;

  protected:
    double zh = 0.0, sampling = 1.0;
    bool exact_filter = false;
    std::vector<Fiber> xfibers, yfibers;
};

}  // namespace ocl
