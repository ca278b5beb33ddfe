// Shell-pair structure-of-arrays built on the device, one thread per pair.
//
// For pair i = (A, B) the thread computes every primitive pair's
// K = sqrt(2) pi^{5/4} c_a c_b exp(-ab/p |AB|^2) / p and bound |K|/sqrt(p),
// orders the primitive pairs by bound descending (stable), counts those whose
// bound times kref reaches PRIM_EPS (kpn: the only ones the ERI kernels
// visit; a primitive quartet is at most K_ab K_cd / sqrt(p+q) times the
// polynomial factors of the p-type recurrences, so the dropped ones are below
// 1e-16 up to those factors; kref < 0 keeps every primitive), and writes the
// fields P, P - A, K, 1/2p and the bound with stride npad.  Padding pairs and primitives get
// a unit exponent and zero weight.
#include "engine.h"

namespace pf {

namespace {

__global__ void __launch_bounds__(256) build_pairs_kernel(const Shell *__restrict__ sh, const int *__restrict__ list,
                                                          int n, int npad, int kp, double kref, double *prim,
                                                          double *AB, int *sa, int *sb, int *oa, int *ob, int *kpn,
                                                          double *bmax) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npad) return;
    auto pad_prim = [&](int u) {
        double *col = prim + (size_t)u * PF_NFIELD * npad + i;
        for (int f = 0; f < PF_NFIELD; ++f) col[f * npad] = 0.0;
        col[PF_P * npad] = 1.0;
        col[PF_O2P * npad] = 0.5;
    };
    if (i >= n) {
        for (int u = 0; u < kp; ++u) pad_prim(u);
        for (int d = 0; d < 3; ++d) AB[d * npad + i] = 0.0;
        kpn[i] = 0;
        return;
    }
    const double kfac = sqrt(2.0) * pow(M_PI, 1.25);
    const int ia = list[2 * i], ib = list[2 * i + 1];
    const Shell &A = sh[ia], &B = sh[ib];
    sa[i] = ia;
    sb[i] = ib;
    oa[i] = A.bf;
    ob[i] = B.bf;
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double ab = A.xyz[d] - B.xyz[d];
        AB[d * npad + i] = ab;
        r2 += ab * ab;
    }
    double bnd[KMAX * KMAX];
    int ord[KMAX * KMAX];
    int no = 0;
    for (int x = 0; x < A.np; ++x)
        for (int y = 0; y < B.np; ++y) {
            const double a = A.exp[x], b = B.exp[y], p = a + b;
            const double K = kfac * A.coef[x] * B.coef[y] * exp(-a * b / p * r2) / p;
            bnd[no] = fabs(K) / sqrt(p);
            ord[no] = x * KMAX + y;
            ++no;
        }
    // stable selection sort, bound descending
    for (int s = 0; s < no; ++s) {
        int best = s;
        for (int t = s + 1; t < no; ++t)
            if (bnd[t] > bnd[best]) best = t;
        const double bb = bnd[best];
        const int ob_ = ord[best];
        for (int t = best; t > s; --t) { bnd[t] = bnd[t - 1]; ord[t] = ord[t - 1]; }
        bnd[s] = bb;
        ord[s] = ob_;
    }
    int cnt = 0;
    for (int u = 0; u < no; ++u) {
        if (kref < 0.0 || bnd[u] * kref >= PRIM_EPS) ++cnt;
        const int x = ord[u] / KMAX, y = ord[u] % KMAX;
        const double a = A.exp[x], b = B.exp[y], p = a + b;
        double *col = prim + (size_t)u * PF_NFIELD * npad + i;
        col[PF_P * npad] = p;
        for (int d = 0; d < 3; ++d) {
            const double P = (a * A.xyz[d] + b * B.xyz[d]) / p;
            col[(PF_PX + d) * npad] = P;
            col[(PF_AX + d) * npad] = P - A.xyz[d];
        }
        col[PF_K * npad] = kfac * A.coef[x] * B.coef[y] * exp(-a * b / p * r2) / p;
        col[PF_O2P * npad] = 0.5 / p;
        col[PF_BND * npad] = bnd[u];
    }
    for (int u = no; u < kp; ++u) pad_prim(u);
    kpn[i] = cnt;
    bmax[i] = no ? bnd[0] : 0.0;
}

__global__ void to_f32_kernel(size_t n, const double *x, float *y) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = float(x[i]);
}

__global__ void __launch_bounds__(1024) max_reduce_kernel(int n, const double *x, double *out) {
    __shared__ double sh[16];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) m = fmax(m, x[i]);
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < 16; ++w) m = fmax(m, sh[w]);
        *out = m;
    }
}

int round64(int n) { return ((n + 63) / 64) * 64; }

// prim [kp][PF_NFIELD][npad] -> aos [npad][kp][PF_AOS] (padding fields 0)
template <typename R>
__global__ void __launch_bounds__(256) prim_aos_kernel(int64_t total, int npad, int kp, const R *__restrict__ prim,
                                                       R *__restrict__ aos) {
    const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (e >= total) return;
    const int f = (int)(e % PF_AOS);
    const int64_t r = e / PF_AOS;
    const int u = (int)(r % kp), i = (int)(r / kp);
    aos[e] = f < PF_NFIELD ? prim[((int64_t)u * PF_NFIELD + f) * npad + i] : R(0);
}

// Gradient primitive-pair data, primitives in shell order (x over A, y over
// B), one thread per pair slot: the structure-of-arrays sets w (P - A, K
// 2 alpha) and r (P - A, K), read per lane when the pair is a ket, and the
// records rec (GR_ fields), read by scalar loads when it is the bra.
// Padding slots and primitives: p = 1, 1/2p = 1/2, K = 0.
__global__ void __launch_bounds__(256) grad_pairs_kernel(const Shell *__restrict__ sh, const int *__restrict__ sa,
                                                         const int *__restrict__ sb, int n, int npad, int kp,
                                                         double *w, double *r, double *rec) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npad) return;
    const double kfac = sqrt(2.0) * pow(M_PI, 1.25);
    int u = 0;
    if (i < n) {
        const Shell &A = sh[sa[i]], &B = sh[sb[i]];
        double r2 = 0.0;
        for (int d = 0; d < 3; ++d) r2 += (A.xyz[d] - B.xyz[d]) * (A.xyz[d] - B.xyz[d]);
        for (int x = 0; x < A.np; ++x)
            for (int y = 0; y < B.np; ++y, ++u) {
                const double a = A.exp[x], b = B.exp[y], p = a + b;
                const double K = kfac * A.coef[x] * B.coef[y] * exp(-a * b / p * r2) / p;
                const size_t o = (size_t)u * PF_NFIELD * npad + i;
                double *cw = w + o, *cr = r + o;
                double *g = rec + ((size_t)i * kp + u) * PF_AOS;
                cw[PF_P * npad] = cr[PF_P * npad] = g[GR_P] = p;
                cw[PF_O2P * npad] = cr[PF_O2P * npad] = g[GR_O2P] = 0.5 / p;
                for (int d = 0; d < 3; ++d) {
                    const double P = (a * A.xyz[d] + b * B.xyz[d]) / p;
                    cw[(PF_PX + d) * npad] = cr[(PF_PX + d) * npad] = g[GR_PX + d] = P;
                    cw[(PF_AX + d) * npad] = cr[(PF_AX + d) * npad] = g[GR_AX + d] = P - A.xyz[d];
                    g[GR_BX + d] = P - B.xyz[d];
                }
                cw[PF_K * npad] = g[GR_KA] = 2.0 * a * K;
                cr[PF_K * npad] = g[GR_K] = K;
                g[GR_KB] = 2.0 * b * K;
                cw[PF_BND * npad] = cr[PF_BND * npad] = 0.0;
                for (int f = GR_BZ + 1; f < PF_AOS; ++f) g[f] = 0.0;
            }
    }
    for (; u < kp; ++u) {
        const size_t o = (size_t)u * PF_NFIELD * npad + i;
        for (double *c : {w + o, r + o}) {
            for (int f = 0; f < PF_NFIELD; ++f) c[f * npad] = 0.0;
            c[PF_P * npad] = 1.0;
            c[PF_O2P * npad] = 0.5;
        }
        double *g = rec + ((size_t)i * kp + u) * PF_AOS;
        for (int f = 0; f < PF_AOS; ++f) g[f] = 0.0;
        g[GR_P] = 1.0;
        g[GR_O2P] = 0.5;
    }
}

}  // namespace

void fill_grad_pairset(const PairSet &ps, const Shell *d_shells, double *w, double *r, double *rec, hipStream_t st) {
    if (ps.npad == 0) return;
    grad_pairs_kernel<<<(ps.npad + 255) / 256, 256, 0, st>>>(d_shells, ps.d_sa.p, ps.d_sb.p, ps.n, ps.npad, ps.kp, w,
                                                             r, rec);
    PF_HIP(hipGetLastError());
}

void fill_pairset(PairSet &ps, const std::vector<Shell> &sh, const Shell *d_shells,
                  const std::vector<std::pair<int, int>> &list, double kref, double &bound_max, bool want32,
                  hipStream_t st) {
    ps.n = (int)list.size();
    ps.peps = kref < 0.0 ? 0.0 : PRIM_EPS;
    ps.npad = round64(std::max(ps.n, 1));
    int kp = 1;
    for (auto &pr : list) kp = std::max(kp, sh[pr.first].np * sh[pr.second].np);
    ps.kp = kp;
    ps.sa.resize(ps.n);
    ps.sb.resize(ps.n);
    std::vector<int> flat(2 * (size_t)std::max(ps.n, 1), 0);
    for (int i = 0; i < ps.n; ++i) {
        ps.sa[i] = list[i].first;
        ps.sb[i] = list[i].second;
        flat[2 * i] = list[i].first;
        flat[2 * i + 1] = list[i].second;
    }
    DevBuf<int> dlist;
    dlist.upload(flat.data(), flat.size(), st);
    const size_t nprim = (size_t)kp * PF_NFIELD * ps.npad;
    ps.d_sa.alloc(ps.npad); ps.d_sb.alloc(ps.npad); ps.d_oa.alloc(ps.npad); ps.d_ob.alloc(ps.npad);
    ps.d_kpn.alloc(ps.npad);
    ps.d_AB.alloc(3 * (size_t)ps.npad);
    ps.d_prim.alloc(nprim);
    DevBuf<double> bmax(std::max(ps.n, 1)), red(1);
    build_pairs_kernel<<<(ps.npad + 255) / 256, 256, 0, st>>>(d_shells, dlist.p, ps.n, ps.npad, kp, kref,
                                                              ps.d_prim.p, ps.d_AB.p, ps.d_sa.p, ps.d_sb.p,
                                                              ps.d_oa.p, ps.d_ob.p, ps.d_kpn.p, bmax.p);
    PF_HIP(hipGetLastError());
    const int64_t naos = (int64_t)ps.npad * kp * PF_AOS;
    ps.d_aos.alloc((size_t)naos);
    prim_aos_kernel<double><<<(unsigned)((naos + 255) / 256), 256, 0, st>>>(naos, ps.npad, kp, ps.d_prim.p, ps.d_aos.p);
    if (want32) {
        ps.f_AB.alloc(3 * (size_t)ps.npad);
        ps.f_prim.alloc(nprim);
        to_f32_kernel<<<(unsigned)((3 * (size_t)ps.npad + 255) / 256), 256, 0, st>>>(3 * (size_t)ps.npad, ps.d_AB.p,
                                                                                     ps.f_AB.p);
        to_f32_kernel<<<(unsigned)((nprim + 255) / 256), 256, 0, st>>>(nprim, ps.d_prim.p, ps.f_prim.p);
        ps.f_aos.alloc((size_t)naos);
        prim_aos_kernel<float><<<(unsigned)((naos + 255) / 256), 256, 0, st>>>(naos, ps.npad, kp, ps.f_prim.p,
                                                                               ps.f_aos.p);
    }
    double m = 0.0;
    if (ps.n) {
        max_reduce_kernel<<<1, 1024, 0, st>>>(ps.n, bmax.p, red.p);
        PF_HIP(hipMemcpyAsync(&m, red.p, sizeof(double), hipMemcpyDeviceToHost, st));
    }
    PF_HIP(hipStreamSynchronize(st));
    bound_max = std::max(bound_max, m);
}

}  // namespace pf
