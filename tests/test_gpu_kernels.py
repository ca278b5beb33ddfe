"""GPU parity of the individual kernels against the oracle / library routines.

Tolerances: 1e-12 absolute for FP64 integrals,
1e-13 relative for Boys, 1e-11 for dense FP64 linear algebra.
"""
import numpy as np
import pytest

from conftest import load_xyz

pytestmark = pytest.mark.gpu


def test_boys_matches_exact(gpu_lib):
    import oracle
    import pyfacc
    xs = np.concatenate([[0.0, 1e-12, 1e-7, 1e-3], np.linspace(0.01, 45.0, 997), [39.99, 40.0, 40.01, 80.0, 500.0]])
    for nmax in range(0, 7):
        F = pyfacc.boys(nmax, xs)
        ref = np.array([oracle.boys_exact(nmax, x) for x in xs])
        rel = np.abs(F - ref) / np.abs(ref)
        assert rel.max() < 1e-13, (nmax, rel.max(), xs[np.argmax(rel.max(axis=1))])


@pytest.mark.parametrize("m,n,k,ta,tb", [(7, 7, 7, 0, 0), (33, 17, 45, 1, 0), (64, 96, 31, 0, 1),
                                         (100, 100, 100, 1, 1), (448, 448, 448, 0, 0), (5, 448, 3, 0, 1)])
def test_gemm_mfma(gpu_lib, m, n, k, ta, tb):
    import pyfacc
    rng = np.random.default_rng(m * 1000 + n + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    C = pyfacc.gemm(A, B, ta=bool(ta), tb=bool(tb), alpha=0.7, beta=-0.3, C=C0)
    ref = 0.7 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.3 * C0
    assert np.abs(C - ref).max() < 1e-12 * max(1.0, k)


@pytest.mark.parametrize("n", [1, 2, 7, 33, 112])
def test_jacobi_eigh(gpu_lib, n):
    import pyfacc
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    w, V = pyfacc.eigh(A)
    wr = np.linalg.eigh(A)[0]
    assert np.abs(w - wr).max() < 1e-11 * max(1.0, np.abs(wr).max())
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12
    assert np.abs(A @ V - V * w).max() < 1e-11 * max(1.0, np.abs(wr).max())


# 16 x 16 tiles up to n = 1024 (odd n: the guarded scalar loads; 448: one k
# run per wavefront; 1000: several), 32 x 32 tiles above
@pytest.mark.parametrize("n,nocc", [(7, 5), (40, 13), (112, 56), (449, 200), (448, 320), (1000, 380), (1030, 500)])
def test_sp2_projector(gpu_lib, n, nocc):
    import pyfacc
    rng = np.random.default_rng(n + nocc)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.concatenate([np.sort(rng.uniform(-20, -0.3, nocc)), np.sort(rng.uniform(0.3, 5.0, n - nocc))])
    F = (Q * ev) @ Q.T
    P = pyfacc.projector(F, nocc)
    Pref = Q[:, :nocc] @ Q[:, :nocc].T
    assert np.abs(P - Pref).max() < 1e-11


@pytest.mark.parametrize("name,basis", [("h2o_c1.xyz", "sto-3g"), ("oh.xyz", "sto-3g"), ("h2o_c1.xyz", "sto-6g"),
                                        ("h2o_16.xyz", "sto-3g"), ("h2o_16.xyz", "sto-6g")])
def test_one_electron_vs_oracle(gpu_lib, name, basis):
    import oracle
    import pyfacc
    sp, xyz = load_xyz(name)
    S, T, V = pyfacc.one_electron(sp, xyz, basis)
    m = oracle.Molecule(sp, xyz, basis)
    So, To, Vo = m.one_electron()
    for a, b in ((S, So), (T, To), (V, Vo)):
        assert np.abs(a - b).max() < 1e-12


@pytest.mark.parametrize("name,basis", [("h2o_c1.xyz", "sto-3g"), ("oh.xyz", "sto-6g")])
def test_eri_packed_vs_oracle(gpu_lib, name, basis):
    import oracle
    import pyfacc
    sp, xyz = load_xyz(name)
    e = pyfacc.eri_packed(sp, xyz, basis)
    ref = oracle.Molecule(sp, xyz, basis).eri_packed()
    assert e.shape == ref.shape
    assert np.abs(e - ref).max() < 1e-12


def test_eri_packed_water_dimer(gpu_lib):
    import oracle
    import pyfacc
    sp, xyz = load_xyz("h2o_16.xyz")
    sp, xyz = sp[:6], xyz[:6]
    e = pyfacc.eri_packed(sp, xyz)
    ref = oracle.Molecule(sp, xyz).eri_packed()
    assert np.abs(e - ref).max() < 1e-12


def _rand_density(n, seed):
    rng = np.random.default_rng(seed)
    D = rng.standard_normal((n, n)) * 0.3
    return D + D.T


@pytest.mark.parametrize("stored", [False, True])
def test_jk_rhf_vs_oracle(gpu_lib, stored):
    import oracle
    import pyfacc
    sp, xyz = load_xyz("h2o_16.xyz")
    sp, xyz = sp[:9], xyz[:9]
    m = oracle.Molecule(sp, xyz)
    D = _rand_density(m.nbf, 1)
    J, K = pyfacc.jk(sp, xyz, D, stored=stored, schwarz=1e-15)
    Jo, Ko = m.jk_packed(m.eri_packed(), D)
    assert np.abs(J - Jo).max() < 1e-11
    assert np.abs(K - Ko).max() < 1e-11


def test_jk_uhf_and_fp32(gpu_lib):
    import oracle
    import pyfacc
    sp, xyz = load_xyz("h2o_15oh.xyz")
    sp, xyz = sp[:8], xyz[:8]
    m = oracle.Molecule(sp, xyz)
    D = np.stack([_rand_density(m.nbf, 2), _rand_density(m.nbf, 3)])
    J, K = pyfacc.jk(sp, xyz, D, schwarz=1e-15)
    packed = m.eri_packed()
    Jo, Ko = m.jk_packed(packed, D)
    assert np.abs(J - (Jo[0] + Jo[1])).max() < 1e-11
    assert np.abs(K - Ko).max() < 1e-11
    J32, K32 = pyfacc.jk(sp, xyz, D, fp32=True, schwarz=1e-15)
    scale = np.abs(Jo).max()
    assert np.abs(J32 - (Jo[0] + Jo[1])).max() < 2e-5 * scale
    assert np.abs(K32 - Ko).max() < 2e-5 * scale
