// Adjoint sweep kernels, muscl reconstruction.
#include "sc_adjoint_impl.hpp"

namespace sc {
SC_INSTANTIATE_ADJOINT(kMuscl)
}  // namespace sc
