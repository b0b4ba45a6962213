// Sweep kernels instantiated for one interpolation scheme (one translation
// unit per scheme keeps the build parallel).
#include "sc_batch_impl.hpp"
#include "sc_launch.hpp"

namespace sc {
SC_INSTANTIATE_SCHEME(kWeno3)
SC_INSTANTIATE_BATCH(kWeno3)
}  // namespace sc
