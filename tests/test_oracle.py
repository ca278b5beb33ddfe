"""Pins of the CPU oracle (runs without a GPU).

Each check is independent of the oracle's own formulas: mpmath incomplete
gamma for Boys, closed-form s-type integrals, p-type integrals from mpmath
derivatives of those closed forms with respect to the centres
((x - A_x) e^{-a|r-A|^2} = (1/2a) d/dA_x e^{-a|r-A|^2}), library
eigensolvers, and literature energies.
"""
import itertools
import os

import mpmath as mp
import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_xyz

mp.mp.dps = 30


# ---------------------------------------------------------------- Boys
def boys_mp(n, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * n + 1)
    return mp.gammainc(n + mp.mpf(0.5), 0, x) / (2 * x ** (n + mp.mpf(0.5)))


@pytest.mark.parametrize("x", [0.0, 1e-9, 1e-4, 0.3, 1.0, 2.5, 7.0, 15.0, 29.9, 30.1, 39.9, 40.0, 55.0, 120.0])
def test_boys_exact_vs_mpmath(x):
    F = oracle.boys_exact(8, x)
    for n in range(9):
        ref = float(boys_mp(n, x))
        assert abs(F[n] - ref) <= 2e-15 * ref, (n, x)


def test_boys_trapezoid_restates_reference_rule():
    # boys.f90:21-23 small-x branch and :26-40 trapezoid with N = 1000
    assert oracle.boys_trapezoid(2, 1e-7) == pytest.approx(1 / 5 - 1e-7 / 7, abs=1e-18)
    for n, x in [(0, 1.0), (1, 0.5), (3, 12.0)]:
        t = np.linspace(0.0, 1.0, 1001)
        f = t ** (2 * n) * np.exp(-x * t * t)
        ref = (f.sum() - 0.5 * (f[0] + f[-1])) / 1000
        assert oracle.boys_trapezoid(n, x) == pytest.approx(ref, rel=1e-13)
    # its error against the exact function is the documented compat gap
    gap = max(abs(oracle.boys_trapezoid(0, x) - float(boys_mp(0, x))) for x in np.linspace(0.01, 5, 50))
    assert 1e-9 < gap < 1e-7


# ---------------------------------------------------------------- closed forms
A = np.array([0.1, -0.3, 0.25])
B = np.array([-0.4, 0.5, 1.1])
C = np.array([0.7, 0.2, -0.6])
Dc = np.array([-0.2, -0.9, 0.4])
a, b, c, d = 1.3, 0.45, 0.8, 2.1


def ss_overlap(a, b, A, B):
    p = a + b
    r2 = sum((x - y) ** 2 for x, y in zip(A, B))
    return (mp.pi / p) ** mp.mpf(1.5) * mp.exp(-a * b / p * r2)


def ss_kinetic(a, b, A, B):
    p = a + b
    mu = a * b / p
    r2 = sum((x - y) ** 2 for x, y in zip(A, B))
    return mu * (3 - 2 * mu * r2) * ss_overlap(a, b, A, B)


def ss_nuclear(a, b, A, B, Cc, Z=1):
    p = a + b
    P = [(a * x + b * y) / p for x, y in zip(A, B)]
    r2 = sum((x - y) ** 2 for x, y in zip(A, B))
    pc2 = sum((x - y) ** 2 for x, y in zip(P, Cc))
    return -Z * 2 * mp.pi / p * mp.exp(-a * b / p * r2) * boys_mp(0, p * pc2)


def ss_eri(a, b, c, d, A, B, Cc, Dd):
    p, q = a + b, c + d
    P = [(a * x + b * y) / p for x, y in zip(A, B)]
    Q = [(c * x + d * y) / q for x, y in zip(Cc, Dd)]
    kab = mp.exp(-a * b / p * sum((x - y) ** 2 for x, y in zip(A, B)))
    kcd = mp.exp(-c * d / q * sum((x - y) ** 2 for x, y in zip(Cc, Dd)))
    pq2 = sum((x - y) ** 2 for x, y in zip(P, Q))
    return 2 * mp.pi ** mp.mpf(2.5) / (p * q * mp.sqrt(p + q)) * kab * kcd * boys_mp(0, p * q / (p + q) * pq2)


def test_ss_one_electron_closed_forms():
    S, T, V = oracle.prim_1e(0, 0, a, b, A, B, Z=[3], C=[C])
    assert S[0, 0] == pytest.approx(float(ss_overlap(a, b, A, B)), rel=1e-14)
    assert T[0, 0] == pytest.approx(float(ss_kinetic(a, b, A, B)), rel=1e-13)
    assert V[0, 0] == pytest.approx(float(ss_nuclear(a, b, A, B, C, Z=3)), rel=1e-13)


def test_ss_eri_closed_form():
    e = oracle.prim_eri(0, 0, 0, 0, a, b, c, d, A, B, C, Dc)
    assert e[0, 0, 0, 0] == pytest.approx(float(ss_eri(a, b, c, d, A, B, C, Dc)), rel=1e-13)


def _deriv(fun, centres, dirs):
    """d^k/d(centre_i)_dir of fun(centres) by mpmath numerical differentiation."""
    base = [list(map(mp.mpf, x)) for x in centres]
    vars_ = []
    for ci, di in dirs:
        vars_.append((ci, di))

    def f(*xs):
        cc = [list(x) for x in base]
        for (ci, di), v in zip(vars_, xs):
            cc[ci][di] = v
        return fun(cc)

    # a coordinate named k times is one variable differentiated k times
    uniq = []
    for v in vars_:
        if v not in uniq:
            uniq.append(v)
    vars_ = uniq
    x0 = [base[ci][di] for ci, di in vars_]
    orders = tuple(sum(1 for w in dirs if tuple(w) == v) for v in vars_)
    return mp.diff(f, x0, orders)


# p_dir on centre 0 (exponent a): (1/2a) d/dA_dir
@pytest.mark.parametrize("da,db", [(0, None), (1, None), (2, 0), (0, 2), (1, 1)])
def test_p_one_electron_from_derivatives(da, db):
    lb = 0 if db is None else 1
    S, T, V = oracle.prim_1e(1, lb, a, b, A, B, Z=[2], C=[C])
    dirs = [(0, da)] + ([] if db is None else [(1, db)])
    scale = 1 / (2 * a) * (1 if db is None else 1 / (2 * b))
    for mat, fun in ((S, lambda cc: ss_overlap(a, b, cc[0], cc[1])),
                     (T, lambda cc: ss_kinetic(a, b, cc[0], cc[1])),
                     (V, lambda cc: ss_nuclear(a, b, cc[0], cc[1], C, Z=2))):
        ref = float(scale * _deriv(fun, [A, B], dirs))
        val = mat[da, 0 if db is None else db]
        assert val == pytest.approx(ref, rel=1e-10, abs=1e-14)


@pytest.mark.parametrize("ls,dirs", [
    ((1, 0, 0, 0), (0,)), ((1, 0, 1, 0), (2, 1)), ((1, 1, 0, 0), (0, 0)), ((1, 1, 1, 0), (1, 2, 1)),
    ((1, 1, 1, 1), (0, 1, 2, 0)), ((1, 1, 1, 1), (2, 2, 2, 2)),
])
def test_p_eri_from_derivatives(ls, dirs):
    e = oracle.prim_eri(*ls, a, b, c, d, A, B, C, Dc)
    exps = (a, b, c, d)
    idx, var, scale = [], [], mp.mpf(1)
    it = iter(dirs)
    for centre, l in enumerate(ls):
        if l:
            dd = next(it)
            idx.append(dd)
            var.append((centre, dd))
            scale /= 2 * exps[centre]
        else:
            idx.append(0)
    ref = float(scale * _deriv(lambda cc: ss_eri(a, b, c, d, *cc), [A, B, C, Dc], var))
    assert e[tuple(idx)] == pytest.approx(ref, rel=1e-9, abs=1e-14)


# d components from second derivatives of the s closed forms:
# (x - A_x)^2 g = (1/4a^2) d^2g/dA_x^2 + (1/2a) g, (x - A_x)(y - A_y) g = (1/4a^2) d^2g/dA_x dA_y
D_COMPS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]  # xx xy xz yy yz zz


def _d_from_s(fun, centres, ci, e, comp, extra=()):
    i, j = D_COMPS[comp]
    v = _deriv(fun, centres, [(ci, i), (ci, j)] + list(extra)) / (4 * e * e)
    if i == j:
        if extra:
            v += _deriv(fun, centres, list(extra)) / (2 * e)
        else:
            v += fun([list(map(mp.mpf, x)) for x in centres]) / (2 * e)
    return v


@pytest.mark.parametrize("comp", range(6))
def test_d_one_electron_from_derivatives(comp):
    S, T, V = oracle.prim_1e(2, 0, a, b, A, B, Z=[2], C=[C])
    for mat, fun in ((S, lambda cc: ss_overlap(a, b, cc[0], cc[1])),
                     (T, lambda cc: ss_kinetic(a, b, cc[0], cc[1])),
                     (V, lambda cc: ss_nuclear(a, b, cc[0], cc[1], C, Z=2))):
        ref = float(_d_from_s(fun, [A, B], 0, a, comp))
        assert mat[comp, 0] == pytest.approx(ref, rel=1e-9, abs=1e-13)


@pytest.mark.parametrize("comp,pc", [(0, 0), (1, 2), (3, 1), (4, 0), (5, 2)])
def test_d_p_eri_from_derivatives(comp, pc):
    """(d s | p s): a d component on centre A, p on centre C."""
    e = oracle.prim_eri(2, 0, 1, 0, a, b, c, d, A, B, C, Dc)
    fun = lambda cc: ss_eri(a, b, c, d, *cc)
    ref = _d_from_s(fun, [A, B, C, Dc], 0, a, comp, extra=[(2, pc)]) / (2 * c)
    assert e[comp, 0, pc, 0] == pytest.approx(float(ref), rel=1e-8, abs=1e-13)


def test_parse_d_shells():
    """STO-3G bromine: the D line gives one l = 2 shell (3d), in file order."""
    m = oracle.Molecule(["Br", "H"], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.67]]))
    ls = [s["l"] for s in m.shells]
    # element blocks in file order (H before Br), then Br's S, SP, SP, SP, D
    assert ls == [0, 0, 0, 1, 0, 1, 0, 1, 2] and m.nbf == 1 + 4 + 3 * 3 + 6


# ---------------------------------------------------------------- contracted
def test_water_invariants():
    sp, xyz = load_xyz("h2o_c1.xyz")
    m = oracle.Molecule(sp, xyz)
    S, T, V = m.one_electron()
    assert np.allclose(np.diag(S), 1.0, atol=1e-14)  # contraction renormalised
    assert np.allclose(S, S.T) and np.allclose(T, T.T) and np.allclose(V, V.T)
    assert np.linalg.eigvalsh(T).min() > 0
    eri = m.eri_full()
    n = m.nbf
    assert eri.shape == (n,) * 4
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)]:
        assert np.abs(eri - eri.transpose(perm)).max() == 0.0
    diag = np.einsum("ijij->ij", eri)
    assert np.all(eri ** 2 <= np.einsum("ij,kl->ijkl", diag, diag) * (1 + 1e-12) + 1e-30)
    # packed index is the canonical order of rscf.f90:67-88
    assert len(m.eri_packed()) == (n * (n + 1) // 2) * (n * (n + 1) // 2 + 1) // 2


def test_jk_matches_full_tensor_and_direct():
    sp, xyz = load_xyz("oh.xyz")
    m = oracle.Molecule(sp, xyz)
    rng = np.random.default_rng(5)
    D = rng.standard_normal((m.nbf, m.nbf))
    D = D + D.T
    eri = m.eri_full()
    Jr = np.einsum("ijkl,kl->ij", eri, D)
    Kr = np.einsum("ikjl,kl->ij", eri, D)
    J, K = m.jk_packed(m.eri_packed(), D)
    assert np.abs(J - Jr).max() < 1e-12 and np.abs(K - Kr).max() < 1e-12
    J2, K2, _ = m.jk_direct(D, thresh=0.0)
    assert np.abs(J2 - Jr).max() < 1e-12 and np.abs(K2 - Kr).max() < 1e-12


def test_jk_direct_timing_samples_partition_the_build():
    """The CPU-baseline samples (every stride-th shell quartet from a given
    offset) for all offsets of one stride are disjoint and sum to the whole
    build, quartet counts included."""
    sp, xyz = load_xyz("h2o_c1.xyz")
    sp, xyz = sp * 3, np.concatenate([xyz, xyz + [0.0, 0.0, 5.5], xyz + [5.5, 0.0, 0.0]])
    m = oracle.Molecule(sp, xyz)
    rng = np.random.default_rng(9)
    D = rng.standard_normal((m.nbf, m.nbf))
    D = D + D.T
    J, K, done = m.jk_direct(D, thresh=1e-10)
    parts = [m.jk_direct(D, thresh=1e-10, stride=3, phase=ph) for ph in range(3)]
    assert sum(p[2] for p in parts) == done and min(p[2] for p in parts) > 0
    assert np.abs(sum(p[0] for p in parts) - J).max() < 1e-12
    assert np.abs(sum(p[1] for p in parts) - K).max() < 1e-12


def test_jacobi_matches_lapack():
    rng = np.random.default_rng(3)
    M = rng.standard_normal((23, 23))
    M = M + M.T
    import ctypes
    A = np.ascontiguousarray(M.copy())
    w = np.zeros(23)
    V = np.zeros((23, 23))
    p = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    oracle.lib().or_jacobi(23, p(A), p(w), p(V))
    assert np.abs(w - np.linalg.eigvalsh(M)).max() < 1e-12
    assert np.abs(M @ V - V * w).max() < 1e-11


# ---------------------------------------------------------------- SCF
CRAWFORD = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                     [-1.638036840407, 1.136548822547, 0.0]])


def test_crawford_known_answer():
    """T. D. Crawford, programming project #3: E_nuc = 8.002367061810450,
    E_SCF = -74.942079928192 Eh with the 8-digit 1969 STO-3G values."""
    m = oracle.Molecule(["O", "H", "H"], CRAWFORD, basis_path=os.path.join(GOLDEN, "sto-3g-hsp1969.txt"))
    assert m.enuc() == pytest.approx(8.002367061810450, abs=1e-12)
    r = m.scf(5, 5, eps=1e-12, max_iter=200, diis=6)
    assert r["converged"]
    assert abs(r["energy"] - (-74.942079928192)) < 1e-9


def test_h2_szabo_ostlund():
    """Szabo & Ostlund section 3.5.2: H2 STO-3G at R = 1.4 bohr, E = -1.117 Eh."""
    m = oracle.Molecule(["H", "H"], np.array([[0, 0, 0], [0, 0, 1.4]]),
                        basis_path=os.path.join(GOLDEN, "sto-3g-hsp1969.txt"))
    r = m.scf(1, 1, eps=1e-12)
    assert r["energy"] == pytest.approx(-1.1167, abs=1e-4)


def test_scf_invariants_and_plain_roothaan():
    sp, xyz = load_xyz("h2o_c1.xyz")
    m = oracle.Molecule(sp, xyz)
    r = m.scf(5, 5, eps=1e-10, max_iter=200)
    rd = m.scf(5, 5, eps=1e-10, max_iter=200, diis=6)
    assert r["converged"] and rd["converged"]
    assert abs(r["energy"] - rd["energy"]) < 1e-9
    S, _, _ = m.one_electron()
    D = r["D"]
    assert np.trace(D @ S) == pytest.approx(10.0, abs=1e-9)
    assert np.abs(D @ S @ D - 2 * D).max() < 1e-8
    # rigid rotation + translation
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m2 = oracle.Molecule(sp, xyz @ R.T + np.array([1.0, -2.0, 0.5]))
    assert m2.scf(5, 5, eps=1e-10, max_iter=200)["energy"] == pytest.approx(r["energy"], abs=1e-9)


def test_size_consistency():
    sp, xyz = load_xyz("h2o_c1.xyz")
    e1 = oracle.Molecule(sp, xyz).scf(5, 5, eps=1e-11, diis=6)["energy"]
    # 5000 bohr: the dipole-dipole energy mu^2/R^3 (~0.46/R^3) is below 1e-11
    far = np.vstack([xyz, xyz + np.array([0.0, 0.0, 5000.0])])
    e2 = oracle.Molecule(sp + sp, far).scf(10, 10, eps=1e-11, diis=6)["energy"]
    assert e2 == pytest.approx(2 * e1, abs=1e-9)


# ---------------------------------------------------------------- UHF pins
# The oracle's UHF branch (uscf.f90:149-203: G_s = J[Da + Db] - K[D_s],
# E = 1/2 sum (Da + Db) H + 1/2 sum (Da Fa + Db Fb)) held to what the
# mathematics fixes: one-electron systems, the closed-shell limit, a
# two-electron triplet whose determinant the basis fixes, size-consistency
# and the UHF stationarity conditions.
def _one_electron_exact(m):
    """Lowest generalised eigenvalue of (T + V, S) plus E_nuc: the exact
    energy of a one-electron system in the basis (no self-interaction)."""
    import scipy.linalg
    S, T, V = m.one_electron()
    return scipy.linalg.eigh(T + V, S, eigvals_only=True)[0] + m.enuc()


@pytest.mark.parametrize("name,species,xyz", [
    ("H", ["H"], [[0.0, 0.0, 0.0]]),
    ("H2+", ["H", "H"], [[0.0, 0.0, 0.0], [0.0, 0.0, 2.0]]),
    ("He+", ["He"], [[0.0, 0.0, 0.0]]),
    ("H3++", ["H", "H", "H"], [[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [0.6, 1.5, 0.3]]),
    ("Li++", ["Li"], [[0.0, 0.0, 0.0]]),
    ("LiH+++", ["Li", "H"], [[0.0, 0.0, 0.0], [0.4, 0.9, 2.6]]),
])
@pytest.mark.parametrize("basis", ["sto-3g", "sto-6g"])
def test_uhf_one_electron_systems_are_exact(name, species, xyz, basis):
    """nalpha = 1, nbeta = 0: J[Da] = K[Da] for a single occupied orbital, so
    F_a = H and E is the lowest eigenvalue of H c = e S c.  A wrong exchange
    factor or energy weight leaves self-interaction and fails this."""
    m = oracle.Molecule(species, np.array(xyz), basis)
    r = m.scf(1, 0, eps=1e-12, max_iter=200)
    assert r["converged"] and r["D"].shape == (2, m.nbf, m.nbf)
    assert r["energy"] == pytest.approx(_one_electron_exact(m), abs=1e-12), name
    assert np.abs(r["D"][1]).max() == 0.0


def test_forced_uhf_reproduces_rhf():
    """A closed shell through the UHF branch: same energy and Da = Db = D_RHF / 2."""
    sp, xyz = load_xyz("h2o_c1.xyz")
    m = oracle.Molecule(sp, xyz)
    rr = m.scf(5, 5, eps=1e-11, max_iter=200)
    ru = m.scf(5, 5, eps=1e-11, max_iter=200, force_uhf=True)
    assert rr["converged"] and ru["converged"] and ru["D"].shape == (2, m.nbf, m.nbf)
    assert ru["energy"] == pytest.approx(rr["energy"], abs=1e-10)
    assert np.abs(ru["D"][0] - 0.5 * rr["D"]).max() < 1e-8
    assert np.abs(ru["D"][1] - 0.5 * rr["D"]).max() < 1e-8
    assert ru["iterations"] == rr["iterations"]


def test_uhf_triplet_slater_condon():
    """H2 STO-3G with nalpha = 2, nbeta = 0: both alpha orbitals of the
    two-function basis are occupied, so the determinant is fixed and its
    energy is the Slater-Condon sum over any orthonormal basis of the space
    (here Loewdin orbitals), E = h_11 + h_22 + (11|22) - (12|21) + E_nuc."""
    m = oracle.Molecule(["H", "H"], np.array([[0.0, 0.0, 0.0], [0.3, -0.2, 1.45]]))
    r = m.scf(2, 0, eps=1e-12, max_iter=50)
    assert r["converged"]
    S, T, V = m.one_electron()
    w, U = np.linalg.eigh(S)
    C = U @ np.diag(w ** -0.5) @ U.T  # Loewdin orbitals (columns)
    h = C.T @ (T + V) @ C
    g = np.einsum("pqrs,pi,qj,rk,sl->ijkl", m.eri_full(), C, C, C, C)
    e = h[0, 0] + h[1, 1] + g[0, 0, 1, 1] - g[0, 1, 1, 0] + m.enuc()
    assert r["energy"] == pytest.approx(e, abs=1e-12)
    assert np.abs(r["D"][0] - np.linalg.inv(S)).max() < 1e-10


def test_uhf_oh_stationary_idempotent():
    """OH doublet: the oracle's converged spin densities are idempotent with the
    right traces and commute with Fock matrices built independently from the
    full ERI tensor (G_a = J[Da + Db] - K[Da])."""
    sp, xyz = load_xyz("oh.xyz")
    m = oracle.Molecule(sp, xyz)
    r = m.scf(5, 4, eps=1e-11, max_iter=300, diis=6)
    assert r["converged"] and r["D"].shape[0] == 2
    S, T, V = m.one_electron()
    H = T + V
    eri = m.eri_full()
    Da, Db = r["D"]
    J = np.einsum("pqrs,rs->pq", eri, Da + Db)
    for Ds, nocc in ((Da, 5), (Db, 4)):
        F = H + J - np.einsum("prqs,rs->pq", eri, Ds)
        assert np.trace(Ds @ S) == pytest.approx(nocc, abs=1e-10)
        assert np.abs(Ds @ S @ Ds - Ds).max() < 1e-9
        assert np.abs(F @ Ds @ S - S @ Ds @ F).max() < 1e-8


def test_uhf_size_consistency():
    """OH + H2O 5000 bohr apart: E(UHF, doublet) = E_UHF(OH) + E_RHF(H2O)."""
    sp1, x1 = load_xyz("oh.xyz")
    sp2, x2 = load_xyz("h2o_c1.xyz")
    e1 = oracle.Molecule(sp1, x1).scf(5, 4, eps=1e-11, max_iter=300, diis=6)
    e2 = oracle.Molecule(sp2, x2).scf(5, 5, eps=1e-11, max_iter=300, diis=6)
    far = np.vstack([x1, x2 + np.array([0.0, 0.0, 5000.0])])
    e12 = oracle.Molecule(sp1 + sp2, far).scf(10, 9, eps=1e-11, max_iter=300, diis=6)
    assert e1["converged"] and e2["converged"] and e12["converged"]
    assert e12["energy"] == pytest.approx(e1["energy"] + e2["energy"], abs=1e-9)
