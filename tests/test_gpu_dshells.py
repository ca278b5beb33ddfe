"""Cartesian d shells (the STO-nG D shells of Ga-Br) on the generic
McMurchie-Davidson path, against the oracle: one-electron integrals, every
unique ERI, shell-quartet blocks in any orientation, J/K, RHF/UHF energies
and the mixed-precision SCF.  Bars: integrals 1e-12 absolute, energies 1e-9 Eh."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ANG = 1.8897259886


def _pyramid(r, angle_deg, n):
    """n H atoms at distance r (angstrom) around the origin: n = 1 on z, n = 2
    bent in the xy plane, n = 3 pyramidal (C3v), n = 4 tetrahedral."""
    if n == 1:
        return [[0.0, 0.0, r]]
    if n == 2:
        h = np.radians(angle_deg) / 2
        return [[r * np.sin(h), r * np.cos(h), 0.0], [-r * np.sin(h), r * np.cos(h), 0.0]]
    if n == 3:
        s2 = 2.0 / 3.0 * (1.0 - np.cos(np.radians(angle_deg)))
        st, ct = np.sqrt(s2), np.sqrt(1 - s2)
        return [[r * st * np.cos(f), r * st * np.sin(f), r * ct] for f in np.radians([0.0, 120.0, 240.0])]
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(3.0)
    return list(r * v)


MOLS = {  # experimental bond lengths (angstrom) and angles
    "HBr": (["Br", "H"], 1.414, 0.0),
    "H2Se": (["Se", "H", "H"], 1.460, 91.0),
    "AsH3": (["As", "H", "H", "H"], 1.511, 92.1),
    "GeH4": (["Ge", "H", "H", "H", "H"], 1.525, 109.47),
}


def _mol(name):
    sp, r, ang = MOLS[name]
    xyz = np.array([[0.0, 0.0, 0.0]] + _pyramid(r, ang, len(sp) - 1)) * ANG
    # a small tilt so no coordinate axis is special
    c, s = np.cos(0.3), np.sin(0.3)
    R = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]]) @ np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])
    return sp, xyz @ R.T + np.array([0.1, -0.2, 0.3])


@pytest.mark.parametrize("name", list(MOLS))
@pytest.mark.parametrize("basis", ["sto-3g", "sto-6g"])
def test_d_one_electron_vs_oracle(gpu_lib, name, basis):
    import oracle
    import pyfacc
    sp, xyz = _mol(name)
    S, T, V = pyfacc.one_electron(sp, xyz, basis)
    So, To, Vo = oracle.Molecule(sp, xyz, basis).one_electron()
    for M, Mo in ((S, So), (T, To), (V, Vo)):
        assert np.abs(M - Mo).max() < 1e-12 * max(1.0, np.abs(Mo).max())


@pytest.mark.parametrize("name", ["HBr", "AsH3"])
def test_d_every_unique_eri_vs_oracle(gpu_lib, name):
    import oracle
    import pyfacc
    sp, xyz = _mol(name)
    got = pyfacc.eri_packed(sp, xyz)
    ref = oracle.Molecule(sp, xyz).eri_packed()
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() < 1e-12


def test_d_shell_quartet_blocks_any_orientation(gpu_lib):
    """pf_eri_shell_quartets on quartets with the d shell in every position."""
    import oracle
    import pyfacc
    sp, xyz = _mol("H2Se")
    m = oracle.Molecule(sp, xyz)
    L = m.basis.l
    d = int(np.where(L == 2)[0][0])
    p = int(np.where(L == 1)[0][-1])
    s = int(np.where(L == 0)[0][0])
    quartets = [(d, s, s, s), (s, d, p, s), (p, s, d, d), (d, d, d, d), (s, p, s, d), (d, p, d, s), (p, d, p, p)]
    blocks = pyfacc.eri_quartets(sp, xyz, quartets)
    for q, blk in zip(quartets, blocks):
        ref = m.shell_eri(*q)
        assert blk.shape == ref.shape, q
        assert np.abs(blk - ref).max() < 1e-12, q


@pytest.mark.parametrize("name", ["HBr", "GeH4"])
def test_d_jk_vs_oracle(gpu_lib, name):
    import oracle
    import pyfacc
    sp, xyz = _mol(name)
    m = oracle.Molecule(sp, xyz)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((2, m.nbf, m.nbf))
    D = np.ascontiguousarray(A + A.transpose(0, 2, 1)) * 0.1
    Jo, Ko = m.jk_packed(m.eri_packed(), D)
    J, K = pyfacc.jk(sp, xyz, D, schwarz=1e-16)
    assert np.abs(J - (Jo[0] + Jo[1])).max() < 1e-11
    assert np.abs(K - Ko).max() < 1e-11
    J32, K32 = pyfacc.jk(sp, xyz, D, schwarz=1e-16, fp32=True)
    assert np.abs(J32 - J).max() < 2e-5 * np.abs(J).max() and np.abs(K32 - K).max() < 2e-5 * np.abs(K).max()


@pytest.mark.parametrize("name", list(MOLS))
def test_d_rhf_energy_vs_oracle(gpu_lib, name):
    import oracle
    import pyfacc
    sp, xyz = _mol(name)
    ne = int(sum(pyfacc.symbols_to_z(sp)))
    ref = oracle.Molecule(sp, xyz).scf(ne // 2, ne // 2, eps=1e-10, max_iter=300, diis=8)
    assert ref["converged"]
    r = pyfacc.scf(sp, xyz, ne // 2, ne // 2, eps=1e-10, max_iter=300, diis=8, schwarz=1e-14)
    assert r["converged"]
    assert abs(r["energy"] - ref["energy"]) < 1e-9
    mx = pyfacc.scf(sp, xyz, ne // 2, ne // 2, eps=1e-10, max_iter=300, diis=8, precision=1, guess=1)
    assert mx["converged"] and abs(mx["energy"] - r["energy"]) < 1e-8


def test_d_uhf_cation_vs_oracle(gpu_lib):
    """HBr+ doublet (nalpha = 18, nbeta = 17) through the UHF branch."""
    import oracle
    import pyfacc
    sp, xyz = _mol("HBr")
    ref = oracle.Molecule(sp, xyz).scf(18, 17, eps=1e-10, max_iter=300, diis=8)
    assert ref["converged"]
    r = pyfacc.scf(sp, xyz, 18, 17, eps=1e-10, max_iter=300, diis=8, schwarz=1e-14)
    assert r["converged"] and abs(r["energy"] - ref["energy"]) < 1e-9


def test_d_paths_not_built_refuse_cleanly(gpu_lib):
    """Gradients, stored ERIs and the batched engine cover s and p shells only:
    they refuse a basis with d shells with a clear message."""
    import pyfacc
    from pyfacc.grad import scf_gradient
    sp, xyz = _mol("HBr")
    with pytest.raises(pyfacc.PyfaccError, match="s and p shells only"):
        scf_gradient(sp, xyz, 18, 18, eps=1e-9)
    with pytest.raises(pyfacc.PyfaccError, match="s and p shells only"):
        pyfacc.scf(sp, xyz, 18, 18, eri_mode=1)
    with pytest.raises(pyfacc.PyfaccError, match="s and p shells only"):
        pyfacc.batch_scf([(sp, xyz, 18, 18)])
