// Instantiates the ERI/J/K kernels of quartet class (ss|ss); one class per
// translation unit so the classes compile in parallel.
#include "eri_kernel.h"

namespace pf {
template void launch_nk_mode<0, 0, 0, 0, double>(int, int, const JKArgs<double> &, hipStream_t);
template void launch_nk_mode<0, 0, 0, 0, float>(int, int, const JKArgs<float> &, hipStream_t);
}  // namespace pf
