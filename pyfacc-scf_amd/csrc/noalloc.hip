// pf_projector_noalloc: the SP2 spectral projector of pf_projector for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops that forbid hipMalloc).  The caller passes F and P in device
// memory; the purification's iterate, its ping-pong buffer and its trace and
// convergence records live in `scratch`, a file-scope device array of fixed
// size, so the call allocates nothing and never waits on the device.
#include <mutex>
#include <string>

#include "linalg.h"
#include "pf_common.h"

extern "C" {
#include "../../include/pyfacc.h"
}

namespace pf {
void set_last_error(const std::string &msg);

namespace {

constexpr int NOALLOC_NMAX = PF_PROJECTOR_NOALLOC_NMAX;
// X and work (n x n each) and the SP2 state at the largest n
constexpr size_t SCRATCH_DOUBLES = 2 * (size_t)NOALLOC_NMAX * NOALLOC_NMAX + sp2_state_size(NOALLOC_NMAX);

// One copy per GPU, held until the process ends.  Host code never uses this
// name as an address: the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};

// Device address of this GPU's scratch, resolved once per device with that
// device current (each GPU holds its own copy at its own address).
double *scratch_on(int dev) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(std::string("pf_projector_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    return g_scratch[dev];
}

}  // namespace
}  // namespace pf

int pf_projector_noalloc(int n, const double *F, double nocc, double *P, int max_steps, void *stream) {
    using namespace pf;
    set_last_error("");
    try {
        if (n <= 0 || n > NOALLOC_NMAX)
            throw Error("pf_projector_noalloc: n = " + std::to_string(n) + " outside 1.." +
                        std::to_string(NOALLOC_NMAX) + " (the fixed scratch size)");
        if (!F || !P) throw Error("pf_projector_noalloc: null F or P");
        if (max_steps < 1 || max_steps > 200) throw Error("pf_projector_noalloc: max_steps must be 1..200");
        if (!(nocc >= 0.0 && nocc <= n)) throw Error("pf_projector_noalloc: nocc outside 0..n");
        int dev = 0;
        PF_HIP(hipGetDevice(&dev));
        if (dev < 0 || dev >= MAX_DEVICES) throw Error("pf_projector_noalloc: device index out of range");
        double *base = scratch_on(dev);
        const size_t nn = (size_t)n * n;
        if (2 * nn + sp2_state_size(n) > SCRATCH_DOUBLES) throw Error("pf_projector_noalloc: scratch layout");
        double *X = base, *work = base + nn, *state = base + 2 * nn;
        sp2_project_fixed(n, F, nocc, P, X, work, state, max_steps, static_cast<hipStream_t>(stream));
        return PF_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return std::string(e.what()).find("HIP") != std::string::npos ? PF_ERR_DEVICE : PF_ERR_INPUT;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PF_ERR_INTERNAL;
    }
}
