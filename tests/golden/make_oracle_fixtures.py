"""Writes oracle_fixtures.json: the small reference vectors the parity tests
read (test infrastructure; nothing here ships).

* Boys F_n(x), n = 0..5, at 40 arguments, from mpmath's incomplete gamma
  (30 significant digits, independent of the oracle).
* H2O (h2o_c1.xyz) and OH (oh.xyz), STO-3G: S, T, V with renormalised
  contractions and with primitive norms only (the reference's convention),
  V also with the reference's trapezoid Boys rule (compat mode); the 406
  unique H2O ERIs in canonical packed order; J and K of H2O at a seeded
  symmetric density; the RHF H2O and UHF OH energies (DIIS 8, eps 1e-10).
* (H2O)16 (h2o_16.xyz, 112 basis functions) in STO-3G and its first four
  waters, (H2O)4 (28 basis functions), in STO-6G: the RHF energy (stored
  ERIs, no screening, DIIS 8, eps 1e-10) and checksums of J and K at a
  seeded density (sum, Frobenius norm, and the sum weighted by a fixed +-1
  pattern).

Every number is written with repr(), so the JSON holds it exactly.  Run from
the repository root: python tests/golden/make_oracle_fixtures.py (a few
minutes on 8 cores, most of it the (H2O)16 part).  With --h2o16-sto6g it
only adds the full (H2O)16 cluster in STO-6G (two passes over its 6.8e9
primitive quartets: about an hour on 8 cores) to the existing file; with
--h2o64-jk it only adds checksums of J and K of the (H2O)64 cluster
(h2o_64.xyz, 448 basis functions, STO-3G) at a seeded density, one direct
build screened at Q_ab Q_cd < 1e-12 (a few minutes)."""
import json
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle  # noqa: E402
from conftest import load_xyz  # noqa: E402

BOYS_X = [0.0, 1e-12, 1e-8, 1e-6, 2e-6, 1e-4, 1e-3, 0.01, 0.05, 0.1, 0.2, 0.35, 0.5, 0.75, 1.0, 1.5, 2.0, 2.5,
          3.0, 4.0, 5.0, 6.0, 7.5, 9.0, 11.0, 13.0, 15.0, 17.5, 20.0, 23.0, 26.0, 30.0, 35.0, 40.0, 50.0, 65.0,
          80.0, 100.0, 150.0, 300.0]


def boys_mp(n, x):
    mp.mp.dps = 30
    x = mp.mpf(x)
    if x == 0:
        return 1.0 / (2 * n + 1)
    # F_n(x) = gamma(n + 1/2, 0..x) / (2 x^(n + 1/2))
    return float(mp.gammainc(n + mp.mpf(1) / 2, 0, x) / (2 * x ** (n + mp.mpf(1) / 2)))


def seeded_density(n, seed):
    """Symmetric density-like matrix: C C^T for a random n x (n//2) C, scaled."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C = rng.standard_normal((n, max(1, n // 2))) / np.sqrt(n)
    return C @ C.T


def sign_pattern(n):
    i = np.arange(n)
    return np.where(((i[:, None] * 7 + i[None, :] * 3) % 5) < 2, 1.0, -1.0)


def checksums(M):
    return {"sum": float(M.sum()), "fro": float(np.linalg.norm(M)), "signed": float((M * sign_pattern(len(M))).sum())}


def mat(M):
    return [float(v) for v in np.asarray(M).reshape(-1)]


def small(name, na, nb):
    sp, xyz = load_xyz(name)
    out = {"species": sp, "xyz_bohr": mat(xyz), "nalpha": na, "nbeta": nb}
    m = oracle.Molecule(sp, xyz)
    S, T, V = m.one_electron()
    out.update(S=mat(S), T=mat(T), V=mat(V), nbf=m.nbf)
    oracle.set_compat_boys(True)
    try:
        out["V_compat"] = mat(m.one_electron()[2])
    finally:
        oracle.set_compat_boys(False)
    mp_ = oracle.Molecule(sp, xyz, renorm=False)
    S0, T0, V0 = mp_.one_electron()
    out.update(S_prim_norm=mat(S0), T_prim_norm=mat(T0), V_prim_norm=mat(V0))
    r = m.scf(na, nb, eps=1e-10, max_iter=200, diis=8)
    assert r["converged"]
    out.update(energy=r["energy"], e_nuc=r["e_nuc"])
    return m, out


def cluster_entry(nw, basis):
    sp, xyz = load_xyz("h2o_16.xyz")
    sp, xyz = sp[:3 * nw], xyz[:3 * nw]
    m = oracle.Molecule(sp, xyz, basis)
    r = m.scf(5 * nw, 5 * nw, eps=1e-10, max_iter=200, diis=8)
    assert r["converged"]
    D = seeded_density(m.nbf, 16)
    J, K, _ = m.jk_direct(D)
    return {"basis": basis, "waters": nw, "nbf": m.nbf, "energy": r["energy"], "iterations": r["iterations"],
            "density_seed": 16, "D": checksums(D), "J": checksums(J), "K": checksums(K)}


def main():
    path = os.path.join(HERE, "oracle_fixtures.json")
    if "--h2o64-jk" in sys.argv:
        with open(path) as fh:
            fx = json.load(fh)
        sp, xyz = load_xyz("h2o_64.xyz")
        m = oracle.Molecule(sp, xyz)
        D = seeded_density(m.nbf, 64)
        J, K, done = m.jk_direct(D, thresh=1e-12)
        fx["h2o64_sto3g_jk"] = {"basis": "sto-3g", "nbf": m.nbf, "density_seed": 64, "schwarz": 1e-12,
                                "quartets": int(done), "D": checksums(D), "J": checksums(J), "K": checksums(K)}
        print("h2o64_sto3g_jk", fx["h2o64_sto3g_jk"]["J"], flush=True)
        with open(path, "w") as fh:
            json.dump(fx, fh)
        return
    if "--h2o16-sto6g" in sys.argv:
        with open(path) as fh:
            fx = json.load(fh)
        fx["h2o16_sto6g"] = cluster_entry(16, "sto-6g")
        print("h2o16_sto6g", fx["h2o16_sto6g"]["energy"], flush=True)
        with open(path, "w") as fh:
            json.dump(fx, fh)
        return
    fx = {"boys": {"x": BOYS_X, "nmax": 5, "F": [[boys_mp(n, x) for n in range(6)] for x in BOYS_X]}}
    m, h2o = small("h2o_c1.xyz", 5, 5)
    h2o["eri_packed"] = mat(m.eri_packed())
    D = seeded_density(m.nbf, 7)
    h2o["density_seed"] = 7
    J, K = m.jk_packed(m.eri_packed(), D)
    h2o.update(D=mat(D), J=mat(J), K=mat(K))
    fx["h2o"] = h2o
    fx["oh"] = small("oh.xyz", 5, 4)[1]
    for key, nw, basis in (("h2o16_sto3g", 16, "sto-3g"), ("h2o4_sto6g", 4, "sto-6g")):
        fx[key] = cluster_entry(nw, basis)
        print(key, fx[key]["energy"], flush=True)
    with open(path, "w") as fh:
        json.dump(fx, fh)


if __name__ == "__main__":
    main()
