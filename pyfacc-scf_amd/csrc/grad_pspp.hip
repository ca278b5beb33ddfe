// Instantiates the two-electron gradient kernel of quartet class (ps|pp).
#include "grad_kernel.h"

namespace pf {
template void launch_grad2e<1, 0, 1, 1>(int, const GradArgs &, int *, int, hipStream_t);
}  // namespace pf
