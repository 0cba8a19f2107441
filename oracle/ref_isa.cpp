// ORACLE — test infrastructure only.
//
// The reference's own embree::rcp and IntegerSampleStatistics
// (common/math/emath.h:61-88, data/SampleStatistics.h:195-342, used where
// they lie under the reference tree) compiled with the flags of one device
// type (openpgl/CMakeLists.txt:125-131): built once with the CPU_4 flags
// (_ref/libpglref_w4.so) and once with the CPU_16 flags
// (_ref/libpglref_w16.so), each in its own library so the differently
// compiled inline functions never meet.  Pins the oracle's per-width
// reciprocal and integer statistics.
#include <cstddef>
#include <cstdint>
#include <iostream>
#include <sstream>
#include <string>

#include "data/SampleStatistics.h"

using namespace openpgl;

extern "C" {

float ref_isa_rcp(float x) { return embree::rcp(x); }

void ref_isa_integer_stats(const float* pos, size_t n, const float* lo, const float* hi, int64_t* raw) {
  BBox b(Vector3(lo[0], lo[1], lo[2]), Vector3(hi[0], hi[1], hi[2]));
  IntegerSampleStatistics is(b);
  for (size_t i = 0; i < n; i++) is.addSample(Point3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]));
  for (int j = 0; j < 3; j++) {
    raw[j] = is.mean[j];
    raw[3 + j] = is.variance[j];
    raw[7 + j] = is.sampleBounds.lower[j];
    raw[10 + j] = is.sampleBounds.upper[j];
  }
  raw[6] = is.numSamples;
}

}  // extern "C"
