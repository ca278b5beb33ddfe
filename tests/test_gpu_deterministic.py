"""Deterministic J/K accumulation (pf_options.deterministic): every addition
into J', K' is an int64 fixed-point add, so repeated SCFs return bit-identical
energies and densities (FP64 atomics make them differ by ~1e-11), in RHF,
UHF and mixed precision and on the fallback paths of large systems; the
result stays within the 1e-9 Eh bar of the FP64-atomic build and of the
oracle."""
import numpy as np
import pytest

from conftest import load_xyz

pytestmark = pytest.mark.gpu

OPTS = dict(eps=1e-9, max_iter=100, diis=8, schwarz=1e-11, guess=1)


def _twice(sp, xyz, na, nb, **kw):
    import pyfacc
    r1 = pyfacc.scf(sp, xyz, na, nb, want_density=True, deterministic=1, **dict(OPTS, **kw))
    r2 = pyfacc.scf(sp, xyz, na, nb, want_density=True, deterministic=1, **dict(OPTS, **kw))
    assert r1["converged"] and r2["converged"]
    return r1, r2


@pytest.mark.parametrize("name,na,nb,kw", [("h2o_16.xyz", 80, 80, {}), ("h2o_15oh.xyz", 80, 79, {}),
                                           ("h2o_16.xyz", 80, 80, {"precision": 1}),
                                           ("h2o_16.xyz", 80, 80, {"jk_fallback": 3})])
def test_repeated_scf_bit_identical(gpu_lib, name, na, nb, kw):
    import pyfacc
    sp, xyz = load_xyz(name)
    r1, r2 = _twice(sp, xyz, na, nb, **kw)
    assert r1["energy"] == r2["energy"] and r1["iterations"] == r2["iterations"]
    assert np.array_equal(r1["D"], r2["D"])
    ref = pyfacc.scf(sp, xyz, na, nb, **dict(OPTS, **kw))
    assert abs(r1["energy"] - ref["energy"]) < 1e-9


@pytest.mark.parametrize("name,na,nb", [("h2o_c1.xyz", 5, 5), ("oh.xyz", 5, 4)])
def test_deterministic_matches_oracle(gpu_lib, name, na, nb):
    import oracle
    import pyfacc
    sp, xyz = load_xyz(name)
    ref = oracle.Molecule(sp, xyz).scf(na, nb, eps=1e-10, max_iter=150, diis=8)
    r = pyfacc.scf(sp, xyz, na, nb, eps=1e-10, max_iter=150, diis=8, schwarz=1e-14, deterministic=1)
    assert ref["converged"] and r["converged"]
    assert abs(r["energy"] - ref["energy"]) < 1e-9


def test_deterministic_jk_fallback_paths_agree(gpu_lib):
    """The device-atomic paths round other partial sums than the LDS-row and
    J'_cd-buffer paths: close, and each one repeatable."""
    sp, xyz = load_xyz("h2o_16.xyz")
    a, _ = _twice(sp, xyz, 80, 80)
    b, _ = _twice(sp, xyz, 80, 80, jk_fallback=3)
    assert abs(a["energy"] - b["energy"]) < 1e-9
    assert np.abs(a["D"] - b["D"]).max() < 1e-8


def test_deterministic_refused_where_not_built(gpu_lib):
    import pyfacc
    sp, xyz = load_xyz("h2o_c1.xyz")
    with pytest.raises(pyfacc.PyfaccError, match="direct builds only"):
        pyfacc.scf(sp, xyz, 5, 5, eri_mode=1, deterministic=1)
    with pytest.raises(pyfacc.PyfaccError, match="batched engine"):
        pyfacc.batch_scf([(sp, xyz, 5, 5)], deterministic=1)
    ang = 1.8897259886
    with pytest.raises(pyfacc.PyfaccError, match="s and p shells only"):
        pyfacc.scf(["Br", "H"], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.414 * ang]]), 18, 18, deterministic=1)
