"""Parity on the seeded (H2O)16 cluster (112 basis functions).

* all 20 024 956 unique ERIs (STO-3G) from the GPU against the oracle, 1e-12;
* GPU J/K (direct and stored) against the oracle's J/K from its packed ERIs;
* RHF (H2O)16 and UHF (H2O)15OH energies against the oracle, 1e-9 Eh;
* STO-6G (contraction depth 6) on a 4-water cut of the same cluster.
The oracle runs on the host cores (OpenMP).
"""
import numpy as np
import pytest

from conftest import load_xyz

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    import oracle
    sp, xyz = load_xyz("h2o_16.xyz")
    m = oracle.Molecule(sp, xyz)
    return sp, xyz, m, m.eri_packed()


def test_c2_all_unique_eris(gpu_lib, c2):
    import pyfacc
    sp, xyz, m, packed = c2
    e = pyfacc.eri_packed(sp, xyz)
    assert e.shape == packed.shape == (20024956,)
    assert np.abs(e - packed).max() < 1e-12


@pytest.mark.parametrize("stored", [False, True])
def test_c2_jk(gpu_lib, c2, stored):
    import pyfacc
    sp, xyz, m, packed = c2
    rng = np.random.default_rng(16)
    C = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))[0][:, :80]
    D = 2 * C @ C.T
    J, K = pyfacc.jk(sp, xyz, D, stored=stored, schwarz=1e-15)
    Jo, Ko = m.jk_packed(packed, D)
    assert np.abs(J - Jo).max() < 1e-10
    assert np.abs(K - Ko).max() < 1e-10


def test_c2_rhf_energy(gpu_lib, c2):
    import pyfacc
    sp, xyz, m, _ = c2
    ref = m.scf(80, 80, eps=1e-9, max_iter=100, diis=8)
    r = pyfacc.scf(sp, xyz, 80, 80, eps=1e-9, max_iter=100, diis=8, schwarz=1e-13)
    assert ref["converged"] and r["converged"]
    assert abs(r["energy"] - ref["energy"]) < 1e-9
    rs = pyfacc.scf(sp, xyz, 80, 80, eps=1e-9, max_iter=100, diis=8, schwarz=1e-13, eri_mode=1)
    assert abs(rs["energy"] - ref["energy"]) < 1e-9


def test_c2_uhf_energy(gpu_lib):
    import oracle
    import pyfacc
    sp, xyz = load_xyz("h2o_15oh.xyz")
    m = oracle.Molecule(sp, xyz)
    ref = m.scf(80, 79, eps=1e-9, max_iter=150, diis=8)
    r = pyfacc.scf(sp, xyz, 80, 79, eps=1e-9, max_iter=150, diis=8, schwarz=1e-13)
    assert ref["converged"] and r["converged"]
    assert abs(r["energy"] - ref["energy"]) < 1e-9


def test_c2_sto6g_cut(gpu_lib):
    import oracle
    import pyfacc
    sp, xyz = load_xyz("h2o_16.xyz")
    sp, xyz = sp[:12], xyz[:12]
    m = oracle.Molecule(sp, xyz, "sto-6g")
    e = pyfacc.eri_packed(sp, xyz, "sto-6g")
    assert np.abs(e - m.eri_packed()).max() < 1e-12
    ref = m.scf(20, 20, eps=1e-10, max_iter=100, diis=8)
    r = pyfacc.scf(sp, xyz, 20, 20, basis="sto-6g", eps=1e-10, max_iter=100, diis=8, schwarz=1e-13)
    assert abs(r["energy"] - ref["energy"]) < 1e-9
