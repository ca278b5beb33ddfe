"""Host-side logic: the drop-in Python API, basis parsing, fixtures, C-ABI exports."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import DATA, ROOT, load_xyz


def test_symbols_and_spin_rules():
    from pyfacc import helper_functions as hf
    assert hf.symbols_to_z(["H", "O", "Xe"]) == [1, 8, 54]
    with pytest.raises(KeyError):
        hf.symbols_to_z(["Og"])
    assert hf.compute_spin_electrons(["O", "H", "H"]) == (5, 5, 1)
    assert hf.compute_spin_electrons(["O", "H"]) == (5, 4, 2)
    assert hf.compute_spin_electrons(["O", "O"]) == (9, 7, 3)   # O2 special case
    # N2 is special-cased to a quartet (helper_functions.py:58-61), which is
    # inconsistent with 14 electrons, so the reference raises here (:82-83)
    with pytest.raises(ValueError):
        hf.compute_spin_electrons(["N", "N"])
    assert hf.compute_spin_electrons(["O", "H"], charge=-1) == (5, 5, 1)
    assert hf.guess_multiplicity(["H"]) == 2


def test_read_xyz_units_layout_and_errors(tmp_path, monkeypatch):
    from pyfacc import helper_functions as hf
    n, xyz, sp = hf.read_xyz("h2o.xyz")
    assert n == 3 and list(sp) == ["O", "H", "H"]
    assert xyz.shape == (3, 3)
    assert xyz[2, 0] == pytest.approx(0.116321 * 1.8897259886)
    # memory order of the (3, natoms) array is x1 y1 z1 x2 ... (what the C-ABI reads)
    flat = np.frombuffer(np.array(xyz, dtype=np.float64).tobytes(order="A"), dtype=np.float64)
    assert flat[:3] == pytest.approx(xyz[:, 0])
    # CWD-relative ./data/molecules takes precedence, like the reference
    d = tmp_path / "data" / "molecules"
    d.mkdir(parents=True)
    (d / "h2o.xyz").write_text("1\nx\nH 0 0 1.0\n")
    monkeypatch.chdir(tmp_path)
    assert hf.read_xyz("h2o.xyz")[0] == 1
    (d / "bad.xyz").write_text("1\nx\nH 0 0\n")
    with pytest.raises(ValueError):
        hf.read_xyz("bad.xyz")


def test_cluster_fixtures_are_reproducible():
    from pyfacc.clusters import water_cluster, batched_instance
    for name, grid, seed, oh in [("h2o_16.xyz", (4, 2, 2), 16, False), ("h2o_15oh.xyz", (4, 2, 2), 16, True),
                                 ("h2o_32.xyz", (4, 4, 2), 32, False), ("h2o_64.xyz", (4, 4, 4), 64, False)]:
        sp, x = water_cluster(grid, seed, oh_first=oh)
        sp_f, x_f = load_xyz(name)
        assert sp == sp_f
        assert np.abs(x * 1.8897259886 - x_f).max() < 1e-9
    sp, x, k, oh = batched_instance(7)
    assert 1 <= k <= 12 and len(sp) == 3 * k - (1 if oh else 0)
    # O-H bond lengths of the rigid monomer survive rotation and jitter
    sp, x = water_cluster((4, 4, 4), 64)
    for m in range(64):
        o, h1, h2 = x[3 * m:3 * m + 3]
        assert np.linalg.norm(h1 - o) == pytest.approx(np.hypot(0.757, 0.586))


@pytest.mark.parametrize("basis", ["sto-2g", "sto-3g", "sto-4g", "sto-5g", "sto-6g"])
def test_library_parser_matches_oracle_parser(basis):
    """The C++ G94 reader gives the shells of the oracle's restatement of basis.f90."""
    import oracle
    import pyfacc
    sp = ["O", "H", "C", "N", "H", "O"]
    xyz = np.arange(18, dtype=float).reshape(6, 3) * 0.7
    info = pyfacc.basis_info(sp, xyz, basis)
    m = oracle.Molecule(sp, xyz, basis)
    assert info["nshell"] == m.basis.nsh and info["nbf"] == m.nbf
    assert list(info["l"]) == [s["l"] for s in m.shells]
    assert list(info["atom"]) == [s["atom"] for s in m.shells]
    assert np.array_equal(info["exps"].reshape(-1), m.basis.alpha)
    assert np.abs(info["coefs"].reshape(-1) - m.basis.coef).max() < 1e-15
    # reference shell order: element blocks in file order (H before C, N, O)
    assert [sp[a] for a in info["atom"][:2]] == ["H", "H"]


@pytest.mark.parametrize("species", [["Zn"], ["Br", "H"], ["As", "H", "H", "H"], ["Se", "H", "H"]])
def test_parser_reads_d_shells_like_the_oracle(species):
    """D lines give Cartesian d shells (6 functions): same shell order,
    exponents and normalised coefficients as the oracle's reader."""
    import oracle
    import pyfacc
    xyz = np.arange(3 * len(species), dtype=float).reshape(-1, 3)
    info = pyfacc.basis_info(species, xyz, "sto-3g")
    m = oracle.Molecule(species, xyz)
    assert 2 in list(info["l"])
    assert list(info["l"]) == [s["l"] for s in m.shells] and info["nbf"] == m.nbf
    assert np.array_equal(info["exps"].reshape(-1), m.basis.alpha)
    assert np.abs(info["coefs"].reshape(-1) - m.basis.coef).max() < 1e-13 * np.abs(m.basis.coef).max()


def test_parser_rejects_f_shells_and_missing_elements(tmp_path):
    import pyfacc
    bad = tmp_path / "f-basis.txt"
    bad.write_text("****\nH     0\nF    1   1.00\n      1.0000000              1.0000000\n****\n")
    with pytest.raises(pyfacc.PyfaccError, match="not supported"):
        pyfacc.basis_info(["H"], np.zeros((1, 3)), str(bad))
    with pytest.raises(pyfacc.PyfaccError, match="not found"):
        pyfacc.basis_info(["H"], np.zeros((1, 3)), "no-such-basis")


def test_every_declared_symbol_is_exported():
    import pyfacc
    header = open(os.path.join(ROOT, "include", "pyfacc.h")).read()
    names = re.findall(r"^\s*(?:void|int|unsigned long long|const char \*)\s*\*?\s*(\w+)\s*\(", header,
                       flags=re.M)
    assert "run_scf_c" in names and "run_scf_interface_py" in names and len(names) >= 14
    lib = ctypes.CDLL(pyfacc.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_no_gpu_paths_fail_loudly():
    import pyfacc
    if pyfacc.device_count() > 0:
        pytest.skip("GPU visible")
    lib = ctypes.CDLL(pyfacc.LIB_PATH)
    lib.pf_last_error.restype = ctypes.c_char_p
    out = ctypes.c_double(0.0)
    z = (ctypes.c_int32 * 1)(1)
    xyz = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    lib.run_scf_c(1, 1, 0, 1, 0, b"H ", xyz, ctypes.c_double(1e-8), 10, ctypes.byref(out), b"sto-3g", z)
    assert out.value != out.value
    assert b"GPU" in lib.pf_last_error()
    with pytest.raises(pyfacc.PyfaccError):
        pyfacc.one_electron(["H", "H"], np.array([[0, 0, 0], [0, 0, 1.4]]))


def test_flop_model_counts():
    from pyfacc import flops
    # (ss|ss): setup + Boys(M=0) + 1 prefactor product + 1 accumulate
    assert flops.per_primitive(0, 0, 0, 0) == 16 + (4 + 3 * 6) + 1 + 1
    assert flops.per_primitive(1, 1, 1, 1) > flops.per_primitive(1, 1, 1, 0) > flops.per_primitive(1, 0, 0, 0)
    assert flops.fock_flops([1, 0, 0, 0, 0, 0], kprim_pair=9) == 81 * flops.per_primitive(0, 0, 0, 0) + \
        flops.per_contracted(0, 0, 0, 0)


def test_batch_rejects_single_molecule_accelerators():
    """The batched engine has no level shift or orthonormal DIIS error; asking
    for them is an input error (raised before any device work)."""
    import pyfacc
    mol = (["H", "H"], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.4]]), 1, 1)
    for bad in ({"level_shift": 0.5}, {"diis_basis": 1}):
        with pytest.raises(pyfacc.PyfaccError, match="batched engine"):
            pyfacc.batch_scf([mol], **bad)


def test_options_layout_matches_the_library():
    """pf_default_options fills the ctypes mirror of pf_options: every field
    reads back the documented default (a field added on one side only would
    shift those after it), and the struct ends where the last field does."""
    import ctypes
    import pyfacc
    o = pyfacc._native.default_options()
    expect = dict(eps=1e-8, max_iter=100, precision=0, incremental=0, eri_mode=0, diis=8, diag=0, renormalize=1,
                  schwarz=1e-10, mixed_switch=1e-5, verbose=0, schwarz_inc=1e-12, rebuild_every=0,
                  schwarz_fp32=1e-6, guess=0, level_shift=0.0, level_shift_until=0.1, diis_basis=0,
                  mixed_anchor=0.1, profile=0, jk_fallback=0, deterministic=0, mixed_finish=0)
    assert [f for f, _ in pyfacc._native.Options._fields_] == list(expect)
    for k, v in expect.items():
        assert getattr(o, k) == v, k
    last, ltype = pyfacc._native.Options._fields_[-1]
    end = getattr(pyfacc._native.Options, last).offset + ctypes.sizeof(ltype)
    assert ctypes.sizeof(o) == (end + 7) // 8 * 8  # padded to the double alignment
