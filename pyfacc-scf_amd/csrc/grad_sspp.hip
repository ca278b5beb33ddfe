// Instantiates the two-electron gradient kernel of quartet class (ss|pp).
#include "grad_kernel.h"

namespace pf {
template void launch_grad2e<0, 0, 1, 1>(int, const GradArgs &, int *, int, hipStream_t);
}  // namespace pf
