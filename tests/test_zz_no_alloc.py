"""pf_projector_noalloc (device buffers, fixed scratch array, no allocation)
against pf_projector on the same matrices, through the C-ABI: two calls in a
row, the second smaller, so the second runs on the scratch the first left."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _matrix(n, nocc, seed):
    """Symmetric F with a known spectrum: gap 0.5 at the occupation boundary."""
    rng = np.random.default_rng(seed)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    w = np.concatenate([-1.0 - rng.random(nocc), 0.5 + rng.random(n - nocc)])
    F = (Q * w) @ Q.T
    return 0.5 * (F + F.T), Q[:, :nocc]


def _hip():
    """The HIP runtime libpyfacc.so is linked against (the caller's own
    allocations: this test plays a caller that owns its device buffers)."""
    import pyfacc
    pyfacc.load_library()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64.so" in ln}
    # torch, if imported, maps its own copy; libpyfacc.so's is the system one
    path = sorted(paths, key=lambda q: ("torch" in q, q))[0]
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_projector_noalloc_matches_projector(gpu_lib):
    import pyfacc
    lib = pyfacc.load_library()
    lib.pf_projector_noalloc.restype = ctypes.c_int
    lib.pf_projector_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                         ctypes.c_int, ctypes.c_void_p]
    hip = _hip()
    H2D, D2H = 1, 2
    bufs = []

    def dev_alloc(nbytes):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        bufs.append(p)
        return p

    try:
        for n, nocc, seed in ((448, 160, 1), (97, 31, 2)):
            F, C = _matrix(n, nocc, seed)
            ref = pyfacc.projector(F, nocc)
            Fd, Pd = dev_alloc(8 * n * n), dev_alloc(8 * n * n)
            P = np.full((n, n), np.nan)
            assert hip.hipMemcpy(Fd, F.ctypes.data, 8 * n * n, H2D) == 0
            assert hip.hipMemcpy(Pd, P.ctypes.data, 8 * n * n, H2D) == 0
            rc = lib.pf_projector_noalloc(n, Fd, float(nocc), Pd, 100, None)
            assert rc == 0, lib.pf_last_error()
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(P.ctypes.data, Pd, 8 * n * n, D2H) == 0
            print(f"n={n}: max |P_noalloc - P| = {np.abs(P - ref).max():.1e}")
            assert np.abs(P - ref).max() <= 1e-13
            assert np.abs(P - C @ C.T).max() < 1e-10  # the projector itself
        # out-of-range sizes and null buffers are refused before anything is launched
        x = dev_alloc(8)
        assert lib.pf_projector_noalloc(1025, x, 1.0, x, 100, None) == 1
        assert b"scratch" in lib.pf_last_error()
        assert lib.pf_projector_noalloc(4, x, 1.0, None, 100, None) == 1
    finally:
        for p in bufs:
            hip.hipFree(p)
