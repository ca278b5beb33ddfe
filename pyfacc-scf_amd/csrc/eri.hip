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

// ---- merged work plan (all classes of one build); see PlanArgs

namespace {
template <typename R>
__device__ __forceinline__ int plan_class(const PlanArgs<R> &p, int blk, int PlanClass<R>::*b0, int PlanClass<R>::*b1) {
    int q = 0;
    while (q < p.n - 1 && blk >= p.c[q].*b1) ++q;
    return q;
}
}  // namespace

// one thread per bra pair of every class: its ket prefix and item count (as jk_count_kernel)
template <typename R>
__global__ void __launch_bounds__(256) plan_count_kernel(PlanArgs<R> p) {
    const int q = plan_class(p, (int)blockIdx.x, &PlanClass<R>::cb0, &PlanClass<R>::cb1);
    const PlanClass<R> &c = p.c[q];
    const int i = ((int)blockIdx.x - c.cb0) * 256 + threadIdx.x;
    if (i >= c.bra.n) return;
    const R qthr = p.tau / (c.bra.Q[i] * (p.dmax ? *p.dmax : R(1)));
    int lo = 0, hi = c.ket.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (c.ket.Q[mid] >= qthr) lo = mid + 1;
        else hi = mid;
    }
    const int nk = c.same ? min(lo, i + 1) : lo;
    c.nket[i] = nk;
    c.item_off[i] = (nk + c.chunk - 1) / c.chunk;
}

// one workgroup per class: the exclusive scans of jk_scan_kernel
template <typename R>
__global__ void __launch_bounds__(1024) plan_scan_kernel(PlanArgs<R> p) {
    const PlanClass<R> &c = p.c[blockIdx.x];
    __shared__ int part[1024];
    __shared__ long long jpart[1024];
    const int n = c.bra.n, t = threadIdx.x, per = (n + 1023) / 1024;
    const int b0 = min(n, t * per), b1 = min(n, b0 + per);
    int s = 0;
    long long js = 0;
    for (int i = b0; i < b1; ++i) {
        s += c.item_off[i];
        js += c.nket[i];
    }
    part[t] = s;
    jpart[t] = js;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
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
        const int k = c.item_off[i];
        c.item_off[i] = off;
        off += k;
        c.joff[i] = joff_;
        joff_ += c.nket[i];
    }
    if (t == 1023) {
        c.item_off[n] = part[1023];
        c.work[0] = 0;
        c.work[1] = part[1023];
        c.work[2] = jpart[1023] <= c.jcap;
    }
}

// one thread per ket of every class with a J'_cd buffer: the end of its bra range (as jk_jrange_kernel)
template <typename R>
__global__ void __launch_bounds__(256) plan_jrange_kernel(PlanArgs<R> p) {
    const int q = plan_class(p, (int)blockIdx.x, &PlanClass<R>::rb0, &PlanClass<R>::rb1);
    const PlanClass<R> &c = p.c[q];
    const int j = ((int)blockIdx.x - c.rb0) * 256 + threadIdx.x;
    if (j >= c.ket.n || !c.jbuf || !c.work[2]) return;
    int lo = c.same ? j : 0, hi = c.bra.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (c.nket[mid] > j) lo = mid + 1;
        else hi = mid;
    }
    c.jhi[j] = lo;
}

// one thread per (ket, component) of every class, blockIdx.y a tile of JRED_TILE bras (as jcd_reduce_kernel)
template <typename R, bool DET>
__global__ void __launch_bounds__(256) plan_jsum_kernel(PlanArgs<R> p) {
    const int q = plan_class(p, (int)blockIdx.x, &PlanClass<R>::sb0, &PlanClass<R>::sb1);
    const PlanClass<R> &c = p.c[q];
    if (!c.jbuf || !c.work[2]) return;
    const int ncd = c.ncd;
    const long long t = (long long)((int)blockIdx.x - c.sb0) * 256 + threadIdx.x;
    const int j = (int)(t / ncd), k = (int)(t - (long long)j * ncd);
    if (j >= c.ket.n) return;
    const int i0 = max((int)blockIdx.y * JRED_TILE, c.same ? j : 0);
    const int i1 = min(c.jhi[j], (int)(blockIdx.y + 1) * JRED_TILE);
    double s = 0.0;
#pragma unroll 8
    for (int i = i0; i < i1; ++i) s += c.jbuf[(c.joff[i] + j) * ncd + k];
    const int nd = ncd == 9 ? 3 : 1;  // ket components: s s (1), p s (3 x 1), p p (3 x 3)
    if (s != 0.0)
        acc_add<DET>(&p.J[(size_t)(c.ket.oa[j] + k / nd) * p.nbf + c.ket.ob[j] + k % nd], s, DET ? *p.fx : 0.0);
}

template <typename R>
void launch_plan(PlanArgs<R> &p, hipStream_t st) {
    int cb = 0, rb = 0, sb = 0;
    for (int q = 0; q < p.n; ++q) {
        PlanClass<R> &c = p.c[q];
        c.cb0 = cb; cb += (c.bra.n + 255) / 256; c.cb1 = cb;
        c.rb0 = rb; rb += (c.ket.n + 255) / 256; c.rb1 = rb;
        c.sb0 = sb; sb += (int)(((long long)c.ket.n * c.ncd + 255) / 256); c.sb1 = sb;
    }
    if (cb > 0) plan_count_kernel<R><<<cb, 256, 0, st>>>(p);
    plan_scan_kernel<R><<<p.n, 1024, 0, st>>>(p);
    PF_HIP(hipGetLastError());
}

template <typename R>
void launch_jcd_sums(PlanArgs<R> &p, hipStream_t st) {
    int rb = 0, sb = 0, ty = 1;
    bool any = false;
    for (int q = 0; q < p.n; ++q) {
        const PlanClass<R> &c = p.c[q];
        rb = c.rb1; sb = c.sb1;
        if (c.jbuf) any = true;
        ty = std::max(ty, (c.bra.n + JRED_TILE - 1) / JRED_TILE);
    }
    if (!any || rb == 0) return;
    plan_jrange_kernel<R><<<rb, 256, 0, st>>>(p);
    if (sb > 0) {
        const dim3 grid((unsigned)sb, (unsigned)ty);
        if (p.fx) plan_jsum_kernel<R, true><<<grid, 256, 0, st>>>(p);
        else plan_jsum_kernel<R, false><<<grid, 256, 0, st>>>(p);
    }
    PF_HIP(hipGetLastError());
}

template <typename R>
int direct_chunk(int la, int lb, int lc, int ld, int scale) {
    const int key = ((la * 2 + lb) * 2 + lc) * 2 + ld;
    return scale * [&] {
    switch (key) {
        case 0: return jk_chunk<0, 0, 0, 0, R>();
        case 8: return jk_chunk<1, 0, 0, 0, R>();
        case 2: return jk_chunk<0, 0, 1, 0, R>();
        case 10: return jk_chunk<1, 0, 1, 0, R>();
        case 12: return jk_chunk<1, 1, 0, 0, R>();
        case 3: return jk_chunk<0, 0, 1, 1, R>();
        case 14: return jk_chunk<1, 1, 1, 0, R>();
        case 11: return jk_chunk<1, 0, 1, 1, R>();
        case 15: return jk_chunk<1, 1, 1, 1, R>();
        default: throw Error("jk: unsupported quartet class");
    }
    }();
}
template int direct_chunk<double>(int, int, int, int, int);
template int direct_chunk<float>(int, int, int, int, int);

template void launch_plan<double>(PlanArgs<double> &, hipStream_t);
template void launch_plan<float>(PlanArgs<float> &, hipStream_t);
template void launch_jcd_sums<double>(PlanArgs<double> &, hipStream_t);
template void launch_jcd_sums<float>(PlanArgs<float> &, hipStream_t);

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
