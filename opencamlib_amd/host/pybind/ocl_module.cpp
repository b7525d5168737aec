// Python binding of the drop-in host API (module opencamlib_amd.lib._ocl,
// re-exported as opencamlib_amd.ocl).  Class and method names are those of
// the reference's Python module (src/pythonlib/ocl_geometry.cpp,
// ocl_cutters.cpp:44-89, ocl_dropcutter.cpp:39-81, ocl_algo.cpp:64-103).
// run() releases the GIL while the device works.
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "ocl/ocl.hpp"

namespace py = pybind11;
using namespace ocl;

PYBIND11_MODULE(_ocl, m) {
    m.doc() = "MI355X batch drop-cutter / push-cutter with the OpenCAMLib Python interface";
    py::register_exception<Error>(m, "Error");

    py::class_<Point>(m, "Point")
        .def(py::init<>())
        .def(py::init<double, double, double>())
        .def(py::init<double, double>())
        .def(py::init<const Point&>())
        .def(py::self + py::self)
        .def(py::self - py::self)
        .def(py::self += py::self)
        .def(py::self -= py::self)
        .def(py::self * double())
        .def(double() * py::self)
        .def("norm", &Point::norm)
        .def("xyNorm", &Point::xyNorm)
        .def("normalize", &Point::normalize)
        .def("dot", &Point::dot)
        .def("cross", &Point::cross)
        .def("xRotate", &Point::xRotate)
        .def("yRotate", &Point::yRotate)
        .def("zRotate", &Point::zRotate)
        .def("isRight", &Point::isRight)
        .def("xyDistance", &Point::xyDistance)
        .def("__str__", &Point::str)
        .def_readwrite("x", &Point::x)
        .def_readwrite("y", &Point::y)
        .def_readwrite("z", &Point::z);

    py::enum_<CCType>(m, "CCType")
        .value("NONE", NONE).value("VERTEX", VERTEX).value("VERTEX_CYL", VERTEX_CYL).value("EDGE", EDGE)
        .value("EDGE_SHAFT", EDGE_SHAFT).value("EDGE_HORIZ", EDGE_HORIZ).value("EDGE_CYL", EDGE_CYL)
        .value("EDGE_BALL", EDGE_BALL).value("EDGE_CONE", EDGE_CONE).value("EDGE_CONE_BASE", EDGE_CONE_BASE)
        .value("EDGE_HORIZ_CYL", EDGE_HORIZ_CYL).value("EDGE_HORIZ_TOR", EDGE_HORIZ_TOR)
        .value("EDGE_POS", EDGE_POS).value("EDGE_NEG", EDGE_NEG).value("FACET", FACET)
        .value("FACET_TIP", FACET_TIP).value("FACET_CYL", FACET_CYL).value("ERROR", ERROR);

    py::class_<CCPoint, Point>(m, "CCPoint")
        .def(py::init<>())
        .def(py::init<const CCPoint&>())
        .def(py::init<double, double, double>())
        .def("__str__", &CCPoint::str)
        .def_readwrite("type", &CCPoint::type);

    py::class_<CLPoint, Point>(m, "CLPoint")
        .def(py::init<>())
        .def(py::init<const CLPoint&>())
        .def(py::init<double, double, double>())
        .def(py::init<double, double, double, const CCPoint&>())
        .def("__str__", &CLPoint::str)
        .def("cc", &CLPoint::getCC)
        .def("getCC", &CLPoint::getCC);

    py::class_<Bbox>(m, "Bbox")
        .def("isInside", &Bbox::isInside)
        .def_readonly("maxpt", &Bbox::maxpt)
        .def_readonly("minpt", &Bbox::minpt);

    py::class_<Triangle>(m, "Triangle")
        .def(py::init<Point, Point, Point>())
        .def("getPoints", &Triangle::getPoints)
        .def("__str__", &Triangle::str)
        .def_property_readonly("p", &Triangle::getPoints)
        .def_readonly("n", &Triangle::n);

    py::class_<STLSurf>(m, "STLSurf")
        .def(py::init<>())
        .def("addTriangle", &STLSurf::addTriangle)
        .def("__str__", &STLSurf::str)
        .def("size", &STLSurf::size)
        .def("rotate", &STLSurf::rotate)
        .def("getBounds", &STLSurf::getBounds)
        .def("getTriangles", &STLSurf::getTriangles)
        .def_property_readonly("tris", &STLSurf::getTriangles)
        .def_readonly("bb", &STLSurf::bb);

    py::class_<STLReader>(m, "STLReader").def(py::init<const std::string&, STLSurf&>());

    py::class_<MillingCutter>(m, "MillingCutter")
        .def("getRadius", &MillingCutter::getRadius)
        .def("getLength", &MillingCutter::getLength)
        .def("getDiameter", &MillingCutter::getDiameter)
        .def("__str__", &MillingCutter::str);
    py::class_<CylCutter, MillingCutter>(m, "CylCutter").def(py::init<double, double>());
    py::class_<BallCutter, MillingCutter>(m, "BallCutter").def(py::init<double, double>());
    py::class_<BullCutter, MillingCutter>(m, "BullCutter").def(py::init<double, double, double>());
    py::class_<ConeCutter, MillingCutter>(m, "ConeCutter")
        .def(py::init<double, double, double>(), py::arg("d"), py::arg("angle"), py::arg("l") = 10.0);
    py::class_<CompositeCutter, MillingCutter>(m, "CompositeCutter")
        .def(py::init<>())
        .def("addCutter", &CompositeCutter::addCutter)
        .def("setLength", &CompositeCutter::setLength);
    py::class_<CompCylCutter, CompositeCutter>(m, "CompCylCutter").def(py::init<double, double>());
    py::class_<CompBallCutter, CompositeCutter>(m, "CompBallCutter").def(py::init<double, double>());
    py::class_<CylConeCutter, CompositeCutter>(m, "CylConeCutter").def(py::init<double, double, double>());
    py::class_<BallConeCutter, CompositeCutter>(m, "BallConeCutter").def(py::init<double, double, double>());
    py::class_<BullConeCutter, CompositeCutter>(m, "BullConeCutter").def(py::init<double, double, double, double>());
    py::class_<ConeConeCutter, CompositeCutter>(m, "ConeConeCutter").def(py::init<double, double, double, double>());

    py::class_<Interval>(m, "Interval")
        .def(py::init<>())
        .def(py::init<double, double>())
        .def_readonly("upper", &Interval::upper)
        .def_readonly("lower", &Interval::lower)
        .def_readonly("lower_cc", &Interval::lower_cc)
        .def_readonly("upper_cc", &Interval::upper_cc)
        .def("updateUpper", &Interval::updateUpper)
        .def("updateLower", &Interval::updateLower)
        .def("empty", &Interval::empty)
        .def("__str__", &Interval::str);

    py::class_<Fiber>(m, "Fiber")
        .def(py::init<>())
        .def(py::init<const Point&, const Point&>())
        .def_readonly("p1", &Fiber::p1)
        .def_readonly("p2", &Fiber::p2)
        .def_readonly("dir", &Fiber::dir)
        .def("addInterval", &Fiber::addInterval)
        .def("point", &Fiber::point)
        .def("tval", &Fiber::tval)
        .def("getInts", [](const Fiber& f) { return f.ints; })
        .def("printInts", [](const Fiber& f) {
            for (const Interval& i : f.ints) py::print(i.str());
        })
        .def("__str__", &Fiber::str);

    py::class_<Line>(m, "Line")
        .def(py::init<>())
        .def(py::init<Point, Point>())
        .def(py::init<const Line&>())
        .def_readwrite("p1", &Line::p1)
        .def_readwrite("p2", &Line::p2)
        .def("length2d", &Line::length2d)
        .def("getPoint", &Line::getPoint);
    py::class_<Arc>(m, "Arc")
        .def(py::init<Point, Point, Point, bool>())
        .def(py::init<const Arc&>())
        .def_readonly("p1", &Arc::p1)
        .def_readonly("p2", &Arc::p2)
        .def_readonly("c", &Arc::c)
        .def_readonly("dir", &Arc::dir)
        .def("length2d", &Arc::length2d)
        .def("getPoint", &Arc::getPoint);
    py::class_<Path>(m, "Path")
        .def(py::init<>())
        .def(py::init<const Path&>())
        .def("append", py::overload_cast<const Line&>(&Path::append))
        .def("append", py::overload_cast<const Arc&>(&Path::append))
        .def("getSpans", [](const Path& p) {
            py::list out;
            for (const Span* s : p.span_list) {
                if (s->type() == LineSpanType)
                    out.append(static_cast<const LineSpan*>(s)->line);
                else
                    out.append(static_cast<const ArcSpan*>(s)->arc);
            }
            return out;
        });

    // setSTL copies the surface to the device; setCutter and setPath keep
    // references, so the Python objects are kept alive with the operation.
    py::class_<BatchDropCutter>(m, "BatchDropCutter")
        .def(py::init<>())
        .def("run", &BatchDropCutter::run, py::call_guard<py::gil_scoped_release>())
        .def("getCLPoints", &BatchDropCutter::getCLPoints)
        .def("clearCLPoints", &BatchDropCutter::clearCLPoints)
        .def("setSTL", &BatchDropCutter::setSTL)
        .def("setCutter", &BatchDropCutter::setCutter, py::keep_alive<1, 2>())
        .def("setThreads", &BatchDropCutter::setThreads)
        .def("getThreads", &BatchDropCutter::getThreads)
        .def("appendPoint", &BatchDropCutter::appendPoint)
        .def("getTrianglesUnderCutter", &BatchDropCutter::getTrianglesUnderCutter)
        .def("getCalls", &BatchDropCutter::getCalls)
        .def("getBucketSize", &BatchDropCutter::getBucketSize)
        .def("setBucketSize", &BatchDropCutter::setBucketSize)
        .def("setPrecision", &BatchDropCutter::setPrecision)
        .def("getPrecision", &BatchDropCutter::getPrecision);

    py::class_<PointDropCutter>(m, "PointDropCutter")
        .def(py::init<>())
        .def("run", py::overload_cast<CLPoint&>(&PointDropCutter::run))
        .def("setSTL", &PointDropCutter::setSTL)
        .def("setCutter", &PointDropCutter::setCutter, py::keep_alive<1, 2>())
        .def("getCalls", &PointDropCutter::getCalls)
        .def("setBucketSize", &PointDropCutter::setBucketSize)
        .def("getBucketSize", &PointDropCutter::getBucketSize)
        .def("setPrecision", &PointDropCutter::setPrecision);

    py::class_<BatchPushCutter>(m, "BatchPushCutter")
        .def(py::init<>())
        .def("run", &BatchPushCutter::run, py::call_guard<py::gil_scoped_release>())
        .def("setSTL", &BatchPushCutter::setSTL)
        .def("setCutter", &BatchPushCutter::setCutter, py::keep_alive<1, 2>())
        .def("setThreads", &BatchPushCutter::setThreads)
        .def("getThreads", &BatchPushCutter::getThreads)
        .def("appendFiber", &BatchPushCutter::appendFiber)
        .def("getOverlapTriangles", &BatchPushCutter::getOverlapTriangles)
        .def("getCLPoints", &BatchPushCutter::getCLPoints)
        .def("getFibers", [](BatchPushCutter& b) { return *b.getFibers(); })
        .def("getCalls", &BatchPushCutter::getCalls)
        .def("setBucketSize", &BatchPushCutter::setBucketSize)
        .def("getBucketSize", &BatchPushCutter::getBucketSize)
        .def("setXDirection", &BatchPushCutter::setXDirection)
        .def("setYDirection", &BatchPushCutter::setYDirection)
        .def("reset", &BatchPushCutter::reset)
        .def("setExactFilter", &BatchPushCutter::setExactFilter)
        .def("setPrecision", &BatchPushCutter::setPrecision);

    py::class_<Waterline>(m, "Waterline")
        .def(py::init<>())
        .def("run", &Waterline::run, py::call_guard<py::gil_scoped_release>())
        .def("setSTL", &Waterline::setSTL, py::keep_alive<1, 2>())
        .def("setCutter", &Waterline::setCutter, py::keep_alive<1, 2>())
        .def("setSampling", &Waterline::setSampling)
        .def("getSampling", &Waterline::getSampling)
        .def("setZ", &Waterline::setZ)
        .def("getZ", &Waterline::getZ)
        .def("reset", &Waterline::reset)
        .def("setThreads", &Waterline::setThreads)
        .def("getXFibers", &Waterline::getXFibers)
        .def("getYFibers", &Waterline::getYFibers)
        .def("getLoops", &Waterline::getLoops)
        .def("getCalls", &Waterline::getCalls)
        .def("setExactFilter", &Waterline::setExactFilter)
        .def("setPrecision", &Waterline::setPrecision);

    py::class_<PathDropCutter>(m, "PathDropCutter")
        .def(py::init<>())
        .def("run", &PathDropCutter::run, py::call_guard<py::gil_scoped_release>())
        .def("getCLPoints", &PathDropCutter::getCLPoints)
        .def("setCutter", &PathDropCutter::setCutter, py::keep_alive<1, 2>())
        .def("setSTL", &PathDropCutter::setSTL)
        .def("setSampling", &PathDropCutter::setSampling)
        .def("getSampling", &PathDropCutter::getSampling)
        .def("setPath", &PathDropCutter::setPath, py::keep_alive<1, 2>())
        .def("getZ", &PathDropCutter::getZ)
        .def("setZ", &PathDropCutter::setZ)
        .def("getCalls", &PathDropCutter::getCalls)
        .def("setPrecision", &PathDropCutter::setPrecision);

    py::class_<AdaptivePathDropCutter, PathDropCutter>(m, "AdaptivePathDropCutter")
        .def(py::init<>())
        .def("run", &AdaptivePathDropCutter::run, py::call_guard<py::gil_scoped_release>())
        .def("setMinSampling", &AdaptivePathDropCutter::setMinSampling)
        .def("setCosLimit", &AdaptivePathDropCutter::setCosLimit)
        .def("setChordTolerance", &AdaptivePathDropCutter::setChordTolerance)
        .def("getChordTolerance", &AdaptivePathDropCutter::getChordTolerance)
        .def("getLevels", &AdaptivePathDropCutter::getLevels)
        .def("getSampling", &AdaptivePathDropCutter::getSampling);
}
