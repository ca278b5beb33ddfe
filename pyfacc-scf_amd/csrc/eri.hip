// Dispatch of the per-class ERI/J/K kernels (instantiated in eri_<class>.hip)
// and the Schwarz-bound kernels.
#include "eri_kernel.h"

namespace pf {

extern template void launch_nk_mode<0, 0, 0, 0, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<0, 0, 0, 0, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<1, 0, 0, 0, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<1, 0, 0, 0, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<0, 0, 1, 0, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<0, 0, 1, 0, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<1, 0, 1, 0, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<1, 0, 1, 0, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<1, 1, 0, 0, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<1, 1, 0, 0, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<0, 0, 1, 1, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<0, 0, 1, 1, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<1, 1, 1, 0, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<1, 1, 1, 0, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<1, 0, 1, 1, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<1, 0, 1, 1, float>(int, int, const JKArgs<float> &, hipStream_t);
extern template void launch_nk_mode<1, 1, 1, 1, double>(int, int, const JKArgs<double> &, hipStream_t);
extern template void launch_nk_mode<1, 1, 1, 1, float>(int, int, const JKArgs<float> &, hipStream_t);

__global__ void __launch_bounds__(1024) jk_scan_kernel(int n, int *item_off, const int *nket, long long *joff,
                                                       long long jcap, int *work) {
    __shared__ int part[1024];
    __shared__ long long jpart[1024];
    const int t = threadIdx.x, per = (n + 1023) / 1024;
    const int b0 = min(n, t * per), b1 = min(n, b0 + per);
    int s = 0;
    long long js = 0;
    for (int i = b0; i < b1; ++i) {
        s += item_off[i];
        js += nket[i];
    }
    part[t] = s;
    jpart[t] = js;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // inclusive scans of the per-thread sums
        const int v = t >= o ? part[t - o] : 0;
        const long long jv = t >= o ? jpart[t - o] : 0;
        __syncthreads();
        part[t] += v;
        jpart[t] += jv;
        __syncthreads();
    }
    int off = part[t] - s;
    long long joff_ = jpart[t] - js;
    for (int i = b0; i < b1; ++i) {
        const int c = item_off[i];
        item_off[i] = off;
        off += c;
        joff[i] = joff_;
        joff_ += nket[i];
    }
    if (t == 1023) {
        item_off[n] = part[1023];
        work[0] = 0;
        work[1] = part[1023];
        work[2] = jpart[1023] <= jcap;  // the J'_cd buffer holds this build's ket lists
    }
}

template <typename R>
void launch_jk_class(int la, int lb, int lc, int ld, int nk, int mode, const JKArgs<R> &args, hipStream_t st) {
    const int key = ((la * 2 + lb) * 2 + lc) * 2 + ld;
    switch (key) {
        case 0: launch_nk_mode<0, 0, 0, 0, R>(nk, mode, args, st); break;  // (ss|ss)
        case 8: launch_nk_mode<1, 0, 0, 0, R>(nk, mode, args, st); break;  // (ps|ss)
        case 2: launch_nk_mode<0, 0, 1, 0, R>(nk, mode, args, st); break;  // (ss|ps)
        case 10: launch_nk_mode<1, 0, 1, 0, R>(nk, mode, args, st); break;  // (ps|ps)
        case 12: launch_nk_mode<1, 1, 0, 0, R>(nk, mode, args, st); break;  // (pp|ss)
        case 3: launch_nk_mode<0, 0, 1, 1, R>(nk, mode, args, st); break;  // (ss|pp)
        case 14: launch_nk_mode<1, 1, 1, 0, R>(nk, mode, args, st); break;  // (pp|ps)
        case 11: launch_nk_mode<1, 0, 1, 1, R>(nk, mode, args, st); break;  // (ps|pp)
        case 15: launch_nk_mode<1, 1, 1, 1, R>(nk, mode, args, st); break;  // (pp|pp)
        default: throw Error("jk: unsupported quartet class");
    }
    PF_HIP(hipGetLastError());
}

template <typename R>
void launch_schwarz(int la, int lb, const PairView<R> &pv, const R *tab, R *Q, hipStream_t st) {
    if (pv.n == 0) return;
    const dim3 grid((pv.n + 63) / 64), block(64);
    if (la == 0 && lb == 0) schwarz_kernel<0, 0, R><<<grid, block, 0, st>>>(pv, tab, Q);
    else if (la == 1 && lb == 0) schwarz_kernel<1, 0, R><<<grid, block, 0, st>>>(pv, tab, Q);
    else if (la == 1 && lb == 1) schwarz_kernel<1, 1, R><<<grid, block, 0, st>>>(pv, tab, Q);
    else throw Error("schwarz: unsupported pair class");
    PF_HIP(hipGetLastError());
}

template void launch_jk_class<double>(int, int, int, int, int, int, const JKArgs<double> &, hipStream_t);
template void launch_jk_class<float>(int, int, int, int, int, int, const JKArgs<float> &, hipStream_t);
template void launch_schwarz<double>(int, int, const PairView<double> &, const double *, double *, hipStream_t);

}  // namespace pf
