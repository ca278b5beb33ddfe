"""The oracle against the committed vectors in tests/golden/oracle_fixtures.json
(written by tests/golden/make_oracle_fixtures.py).

The Boys values come from mpmath, independently of the oracle; the rest are
the oracle's own outputs, kept so that a change to the oracle that moves any
of them shows here (and so that the GPU tests compare against fixed
numbers).  The (H2O)16 entries are checked on the GPU only (test_gpu_golden.py)."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN

FIX = os.path.join(GOLDEN, "oracle_fixtures.json")


@pytest.fixture(scope="module")
def fx():
    with open(FIX) as fh:
        return json.load(fh)


def _mol(entry, renorm=True):
    import oracle
    xyz = np.array(entry["xyz_bohr"]).reshape(-1, 3)
    return oracle.Molecule(entry["species"], xyz, renorm=renorm)


def test_boys_oracle_vs_mpmath_fixture(fx):
    import oracle
    b = fx["boys"]
    for x, ref in zip(b["x"], b["F"]):
        F = oracle.boys_exact(b["nmax"], x)
        assert np.allclose(F, ref, rtol=2e-15, atol=0.0), x


@pytest.mark.parametrize("name", ["h2o", "oh"])
def test_one_electron_fixtures(fx, name):
    import oracle
    e = fx[name]
    n = e["nbf"]
    S, T, V = _mol(e).one_electron()
    for M, key in ((S, "S"), (T, "T"), (V, "V")):
        assert np.abs(M - np.reshape(e[key], (n, n))).max() < 1e-14, key
    S0, T0, V0 = _mol(e, renorm=False).one_electron()
    for M, key in ((S0, "S_prim_norm"), (T0, "T_prim_norm"), (V0, "V_prim_norm")):
        assert np.abs(M - np.reshape(e[key], (n, n))).max() < 1e-14, key
    oracle.set_compat_boys(True)
    try:
        Vc = _mol(e).one_electron()[2]
    finally:
        oracle.set_compat_boys(False)
    assert np.abs(Vc - np.reshape(e["V_compat"], (n, n))).max() < 1e-14
    # the trapezoid rule's error is visible (compat differs from exact)
    assert np.abs(Vc - V).max() > 1e-9


def test_h2o_eri_and_jk_fixtures(fx):
    e = fx["h2o"]
    n = e["nbf"]
    m = _mol(e)
    packed = m.eri_packed()
    assert len(packed) == 406
    assert np.abs(packed - np.array(e["eri_packed"])).max() < 1e-14
    J, K = m.jk_packed(packed, np.reshape(e["D"], (n, n)))
    assert np.abs(J - np.reshape(e["J"], (n, n))).max() < 1e-13
    assert np.abs(K - np.reshape(e["K"], (n, n))).max() < 1e-13


@pytest.mark.parametrize("name", ["h2o", "oh"])
def test_scf_energy_fixtures(fx, name):
    e = fx[name]
    r = _mol(e).scf(e["nalpha"], e["nbeta"], eps=1e-10, max_iter=200, diis=8)
    assert r["converged"]
    assert abs(r["energy"] - e["energy"]) < 1e-10


def test_h2o64_oracle_certificate_is_a_stationary_rhf_density():
    """The committed (H2O)64 certificate (make_h2o64_certificate.py): the
    density it was made from satisfies the oracle's SCF equations (residual
    max |F D S - S D F| from one oracle Fock build), holds 640 electrons and
    is idempotent; its energy functional is then the oracle's SCF energy of the
    headline cluster up to O(residual^2)."""
    import json
    with open(os.path.join(GOLDEN, "h2o64_certificate.json")) as fh:
        c = json.load(fh)
    assert c["oracle_schwarz"] <= 1e-12
    assert c["stationarity_residual_max"] < 1e-7
    assert abs(c["trace_DS"] - 640.0) < 1e-9
    assert c["idempotency_max"] < 1e-10
    assert -4797.5 < c["energy_functional"] < -4797.3
