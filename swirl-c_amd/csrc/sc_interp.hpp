// Stand-alone face interpolation launchers (sc_interp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sc_device.hpp"

namespace sc {

// (neg, pos) = left / right face states of face i-1/2 stored at i along dim
// (0 x, 1 y, 2 z) of the [z][x][y] array with extents N (x, y, z); scheme
// kWeno5, kWeno3 or kMuscl.  Arguments are validated by the caller.
hipError_t launch_face_interp(const float* v, const int N[3], int dim, int scheme, float* neg,
                              float* pos, hipStream_t stream);
// out = a / b or 0 where b == 0 (op 0), a * b (op 1), (1 / b) * a (op 2).
hipError_t launch_elementwise(int op, const float* a, const float* b, float* out, long long n,
                              hipStream_t stream);
// f = value on the plane coordinate[dim] == index of the [z][x][y] array.
hipError_t launch_fill_plane(float* f, const int N[3], int dim, int index, float value,
                             const float* values,
                             hipStream_t stream);
// out[i] = 0.5 * (v[i-1] + v[i]) along dim, v[-1] = 0.
hipError_t launch_linear_interp(const float* v, const int N[3], int dim, float* out,
                                hipStream_t stream);

}  // namespace sc
