// Small device entry points used to test kernels on host buffers.
#include "boys.h"
#include "kernels.h"
#include "linalg.h"

namespace pf {

template <int M>
__global__ void boys_kernel(int npts, const double *x, const double *tab, double *F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npts) return;
    double f[M + 1];
    boys_eval<M, double>(x[i], tab, f);
    for (int m = 0; m <= M; ++m) F[(size_t)i * (M + 1) + m] = f[m];
}

void boys_device(int nmax, int npts, const double *x, double *F, hipStream_t st) {
    const auto tab = boys_table_host();
    DevBuf<double> dt, dx, dF((size_t)npts * (nmax + 1));
    dt.upload(tab.data(), tab.size(), st);
    dx.upload(x, npts, st);
    const int b = (npts + 63) / 64;
    switch (nmax) {
        case 0: boys_kernel<0><<<b, 64, 0, st>>>(npts, dx.p, dt.p, dF.p); break;
        case 1: boys_kernel<1><<<b, 64, 0, st>>>(npts, dx.p, dt.p, dF.p); break;
        case 2: boys_kernel<2><<<b, 64, 0, st>>>(npts, dx.p, dt.p, dF.p); break;
        case 3: boys_kernel<3><<<b, 64, 0, st>>>(npts, dx.p, dt.p, dF.p); break;
        case 4: boys_kernel<4><<<b, 64, 0, st>>>(npts, dx.p, dt.p, dF.p); break;
        case 5: boys_kernel<5><<<b, 64, 0, st>>>(npts, dx.p, dt.p, dF.p); break;
        case 6: boys_kernel<6><<<b, 64, 0, st>>>(npts, dx.p, dt.p, dF.p); break;
        default: throw Error("pf_boys: nmax must be 0..6");
    }
    PF_HIP(hipGetLastError());
    dF.download(F, (size_t)npts * (nmax + 1), st);
    PF_HIP(hipStreamSynchronize(st));
}

}  // namespace pf
