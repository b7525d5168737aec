// Drop-in host API of the MI355X batch drop-cutter / push-cutter: the
// reference's ocl:: classes (geometry, cutters, operations) over the C ABI of
// include/ocl_hip.h.
#pragma once

#include "ocl/cutters.hpp"
#include "ocl/geometry.hpp"
#include "ocl/operations.hpp"
