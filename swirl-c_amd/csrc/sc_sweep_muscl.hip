// Sweep kernels instantiated for one interpolation scheme (one translation
// unit per scheme keeps the build parallel).
#include "sc_launch.hpp"

namespace sc {
SC_INSTANTIATE_SCHEME(kMuscl)
}  // namespace sc
