// Adjoint sweep kernels, weno5 reconstruction.
#include "sc_adjoint_impl.hpp"

namespace sc {
SC_INSTANTIATE_ADJOINT(kWeno5)
}  // namespace sc
