// Adjoint sweep kernels, weno3 reconstruction.
#include "sc_adjoint_impl.hpp"

namespace sc {
SC_INSTANTIATE_ADJOINT(kWeno3)
}  // namespace sc
