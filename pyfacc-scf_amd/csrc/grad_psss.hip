// Instantiates the two-electron gradient kernel of quartet class (ps|ss).
#include "grad_kernel.h"

namespace pf {
template void launch_grad2e<1, 0, 0, 0>(int, const GradArgs &, int *, int, hipStream_t);
}  // namespace pf
