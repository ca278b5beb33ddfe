"""The reference's own larger inputs (data/molecules/gly*.xyz, rna*.xyz) and
molecules with the elements they add (C, N, P).

Small molecules: FP64 energies against the oracle at 1e-9 Eh (same geometry,
basis, DIIS).  The glycine chain and the RNA fragment are beyond the oracle's
reach in a test; there the FP64 and mixed-precision SCFs must both converge
and agree to 1e-6 Eh, and the converged density must satisfy Tr(DS) = N and
D S D = 2 D (RHF) to 1e-8.
"""
import numpy as np
import pytest

from conftest import ANG2BOHR, load_xyz

pytestmark = pytest.mark.gpu

# PH3 (C3v, P-H 1.42 A, HPH 93.5 deg), HCN (linear, C-H 1.065 A, C-N 1.153 A),
# the NH2 (1.024 A, 103.4 deg) and PH2 (1.428 A, 91.7 deg) radicals (UHF),
# geometries in Angstrom
_SIN, _COS = 0.84105, 0.54096


def _bent(x, r, angle):
    a = np.radians(angle) / 2
    return [x, "H", "H"], [[0, 0, 0], [r * np.sin(a), 0, r * np.cos(a)], [-r * np.sin(a), 0, r * np.cos(a)]]


SMALL = {
    "PH3": (["P", "H", "H", "H"],
            [[0, 0, 0]] + [[1.42 * _SIN * np.cos(p), 1.42 * _SIN * np.sin(p), -1.42 * _COS]
                           for p in (0.0, 2 * np.pi / 3, 4 * np.pi / 3)], 9, 9),
    "HCN": (["H", "C", "N"], [[0, 0, -1.065], [0, 0, 0], [0, 0, 1.153]], 7, 7),
    "NH2": (*_bent("N", 1.024, 103.4), 5, 4),
    "PH2": (*_bent("P", 1.428, 91.7), 9, 8),
}


@pytest.mark.parametrize("name", sorted(SMALL))
def test_small_cnp_molecules_vs_oracle(gpu_lib, name):
    import oracle
    import pyfacc
    sp, xyz, na, nb = SMALL[name]
    xyz = np.array(xyz, dtype=float) * ANG2BOHR
    ref = oracle.Molecule(sp, xyz).scf(na, nb, eps=1e-10, max_iter=200, diis=8)
    r = pyfacc.scf(sp, xyz, na, nb, eps=1e-10, max_iter=200, diis=8)
    assert ref["converged"] and r["converged"]
    assert abs(r["energy"] - ref["energy"]) < 1e-9
    m = pyfacc.scf(sp, xyz, na, nb, eps=1e-10, max_iter=200, diis=8, precision=1)
    assert m["converged"] and abs(m["energy"] - ref["energy"]) < 1e-8


def _density_checks(sp, xyz, D, nelec):
    import pyfacc
    S = pyfacc.one_electron(sp, xyz)[0]
    assert abs(np.sum(D * S) - nelec) < 1e-8
    assert np.abs(D @ S @ D - 2 * D).max() < 1e-8


@pytest.mark.parametrize("name,charge,opts", [("gly30.xyz", 0, {}), ("rna1.xyz", 0, {"diis_basis": 1}),
                                              ("rna4.xyz", -16, {})])
def test_reference_molecules_converge_mixed_and_fp64(gpu_lib, name, charge, opts):
    """rna1 neutral (the count the reference's run_scf uses by default) with
    the orthonormal DIIS error; rna4 with one negative charge per phosphate
    (neutral, its gap closes and the SCF does not converge)."""
    import pyfacc
    sp, xyz = load_xyz(name)
    ne = int(sum(pyfacc.helper_functions.periodic_table[s] for s in sp)) - charge
    kw = dict(eps=1e-8, max_iter=150, guess=1, **opts)
    f = pyfacc.scf(sp, xyz, ne // 2, ne // 2, want_density=True, **kw)
    m = pyfacc.scf(sp, xyz, ne // 2, ne // 2, precision=1, **kw)
    assert f["converged"] and m["converged"]
    assert abs(m["energy"] - f["energy"]) < 1e-6
    assert m["fock_builds_fp32"] > 0
    _density_checks(sp, xyz, f["D"], ne)


@pytest.mark.parametrize("opts", [{"level_shift": 0.5}, {"level_shift": 0.3, "level_shift_until": 1e-3},
                                  {"diis_basis": 1}])
def test_level_shift_and_diis_basis_keep_the_energy(gpu_lib, opts):
    """Both change the path of the iteration, not its fixed point."""
    import pyfacc
    sp, xyz = load_xyz("h2o_16.xyz")
    kw = dict(eps=1e-9, max_iter=150, guess=1)
    base = pyfacc.scf(sp, xyz, 80, 80, **kw)
    r = pyfacc.scf(sp, xyz, 80, 80, **kw, **opts)
    assert base["converged"] and r["converged"]
    assert abs(r["energy"] - base["energy"]) < 1e-8


@pytest.mark.parametrize("bad", [{"level_shift": -0.1}, {"diis_basis": 2}])
def test_invalid_accelerator_options_are_rejected(gpu_lib, bad):
    import pyfacc
    sp, xyz = load_xyz("h2o_c1.xyz")
    with pytest.raises(pyfacc.PyfaccError, match="level_shift|diis_basis"):
        pyfacc.scf(sp, xyz, 5, 5, **bad)
