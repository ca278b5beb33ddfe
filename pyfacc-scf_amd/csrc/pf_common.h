// Shared host/device definitions for the MI355X Hartree-Fock engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <memory>
#include <string>
#include <vector>

namespace pf {

constexpr int KMAX = 8;       // max primitives per shell (STO-6G needs 6)
constexpr int LMAX = 2;       // s, p (specialised kernels) and d shells (generic kernels)
constexpr int BOYS_MMAX = 14; // highest Boys order held in the table ((dd|dd) needs 8 + 6 Taylor terms)
constexpr int BOYS_TAYLOR = 7;
constexpr double BOYS_H = 1.0 / 16.0;  // table spacing in x
constexpr double BOYS_XMAX = 40.0;     // beyond this the asymptotic form is used
constexpr int BOYS_NGRID = 642;        // x = 0 .. XMAX + 2h

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define PF_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            throw ::pf::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                              __FILE__ + ":" + std::to_string(__LINE__));                \
    } while (0)

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// One contracted shell.  coef[] already contains the primitive norm of the
// (l,0,0) component (gaussian.f90:80-98) and, by default, the contraction
// renormalisation (Libint convention).
struct Shell {
    int l = 0;
    int np = 0;
    int atom = 0;
    int bf = 0;  // first basis function
    double xyz[3] = {0, 0, 0};
    double exp[KMAX] = {0};
    double coef[KMAX] = {0};
};

struct Molecule {
    int natoms = 0;
    std::vector<int> Z;
    std::vector<double> xyz;  // 3*natoms, bohr, per-atom interleaved
};

struct Basis {
    std::vector<Shell> shells;
    int nbf = 0;
};

// Basis construction (basis.cpp).
Basis build_basis(const std::string &path, const std::vector<std::string> &species, const Molecule &mol,
                  bool renormalise);
// The same in steps, so that a batch parses the file once: the file's lines,
// its element blocks (in file order, shell lines with their rows), and the
// basis of one molecule from the blocks.
struct BasisBlocks {
    struct ShellLine {
        std::string kind;
        int np = 0;
        std::vector<std::vector<double>> rows;
    };
    struct Block {
        std::string label;
        std::vector<ShellLine> shells;
    };
    std::string path;
    std::vector<Block> blocks;
};
std::vector<std::string> read_basis_file(const std::string &path);
BasisBlocks parse_basis_blocks(const std::vector<std::string> &lines, const std::string &path);
Basis build_basis_lines(const std::vector<std::string> &lines, const std::string &path,
                        const std::vector<std::string> &species, const Molecule &mol, bool renormalise);
Basis build_basis_blocks(const BasisBlocks &blocks, const std::vector<std::string> &species, const Molecule &mol,
                         bool renormalise);
std::string resolve_basis_path(const std::string &name);
// The parsed file at path, shared by every call (re-read when it changes).
std::shared_ptr<const BasisBlocks> basis_blocks_cached(const std::string &path);
double nuclear_repulsion(const Molecule &mol);

// A device block from the process-wide pool of large scratch blocks: taken
// with pool_acquire (at least `bytes`, on the current device) and given back
// with pool_release once no pending work uses it.
struct PooledBlock {
    void *p = nullptr;
    size_t bytes = 0;
    int dev = -1;
};
PooledBlock pool_acquire(size_t bytes);
void pool_release(PooledBlock &b);

// Device memory for DevBuf through a per-thread cache of released blocks
// (by device and size class): the library's calls each use one stream and
// synchronise it before returning, so a released block is reused only by the
// same thread, ordered after its last use.  dev_free takes the requested size.
void *dev_alloc(size_t bytes);
void dev_free(void *p, size_t bytes);
// Frees the calling thread's cached blocks, the blocks of ended threads and
// the idle pooled blocks (none of them has pending work).
void release_device_cache();
// Bytes held by the calling thread's cache and the ended threads' list.
size_t cached_device_bytes();
// HIP events from a per-thread pool: release only events whose recorded work
// has completed (after a synchronisation or a successful elapsed-time read).
hipEvent_t event_acquire();
void event_release(hipEvent_t e);

// Device buffer with RAII.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        if (count) p = static_cast<T *>(dev_alloc(count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) dev_free(p, n * sizeof(T));
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        release();
        p = o.p; n = o.n; o.p = nullptr; o.n = 0;
        return *this;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        // synchronous: the host source may be a temporary
        if (count) {
            PF_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
            PF_HIP(hipStreamSynchronize(s));
        }
    }
    void download(T *h, size_t count, hipStream_t s) const {
        if (count) PF_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

}  // namespace pf
