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

namespace pf {

__global__ void copy_kernel(size_t n, const double4 *__restrict__ src, double4 *__restrict__ dst) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Device-to-device copy of `bytes` (rounded down to 32 B) with a plain
// grid-stride kernel; returns the kernel time in ms (HIP events).
double device_copy(size_t bytes, hipStream_t st) {
    const size_t n = bytes / sizeof(double4);
    DevBuf<double4> a(n), b(n);
    PF_HIP(hipMemsetAsync(a.p, 0, n * sizeof(double4), st));
    hipEvent_t e0, e1;
    PF_HIP(hipEventCreate(&e0));
    PF_HIP(hipEventCreate(&e1));
    PF_HIP(hipEventRecord(e0, st));
    copy_kernel<<<2048, 256, 0, st>>>(n, a.p, b.p);
    PF_HIP(hipEventRecord(e1, st));
    PF_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

}  // namespace pf
