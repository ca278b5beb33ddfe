"""Analytic nuclear gradients against central finite differences of the GPU
FP64 energy (h = 1e-4 bohr), translational invariance, and the torch autograd
binding.  Bar: 1e-6 Eh/bohr."""
import numpy as np
import pytest

from conftest import load_xyz

pytestmark = pytest.mark.gpu
H = 1e-4


def _fd(species, xyz, na, nb, coords, **kw):
    import pyfacc
    out = []
    for (atom, x) in coords:
        e = []
        for s in (+1, -1):
            c = xyz.copy()
            c[atom, x] += s * H
            e.append(pyfacc.scf(species, c, na, nb, eps=1e-11, max_iter=200, **kw)["energy"])
        out.append((e[0] - e[1]) / (2 * H))
    return np.array(out)


# The OH radical's beta boundary falls inside its degenerate pi pair; its
# first density comes from eigenvectors (purification cannot form that
# projector), so the SCFs at +h and -h break the symmetry the same way.
@pytest.mark.parametrize("name,na,nb", [("h2o_c1.xyz", 5, 5), ("h2o_c1.xyz", 5, 4), ("oh.xyz", 5, 4)])
def test_gradient_vs_finite_differences_small(gpu_lib, name, na, nb):
    from pyfacc.grad import scf_gradient
    sp, xyz = load_xyz(name)
    e, g, r = scf_gradient(sp, xyz, na, nb, eps=1e-11, max_iter=200)
    assert r["converged"]
    assert np.abs(g.sum(axis=0)).max() < 1e-8  # translational invariance
    coords = [(a, x) for a in range(len(sp)) for x in range(3)]
    fd = _fd(sp, xyz, na, nb, coords)
    assert np.abs(g.reshape(-1) - fd).max() < 1e-6


def test_gradient_water_dimer_sto6g(gpu_lib):
    from pyfacc.grad import scf_gradient
    sp, xyz = load_xyz("h2o_16.xyz")
    sp, xyz = sp[:6], xyz[:6]
    e, g, r = scf_gradient(sp, xyz, 10, 10, basis="sto-6g", eps=1e-11, max_iter=200)
    assert r["converged"]
    coords = [(0, 0), (1, 2), (4, 1), (5, 0)]
    fd = _fd(sp, xyz, 10, 10, coords, basis="sto-6g")
    an = np.array([g[a, x] for a, x in coords])
    assert np.abs(an - fd).max() < 1e-6


def test_gradient_mixed_forward_matches_fp64(gpu_lib):
    from pyfacc.grad import scf_gradient
    sp, xyz = load_xyz("h2o_16.xyz")
    sp, xyz = sp[:12], xyz[:12]
    _, g64, _ = scf_gradient(sp, xyz, 20, 20, eps=1e-10, max_iter=200)
    _, gmx, r = scf_gradient(sp, xyz, 20, 20, eps=1e-10, max_iter=200, precision=1)
    assert r["converged"]
    assert np.abs(g64 - gmx).max() < 1e-6


def test_torch_autograd_binding(gpu_lib):
    torch = pytest.importorskip("torch")
    from pyfacc.grad import scf_energy, scf_gradient
    sp, xyz = load_xyz("h2o_c1.xyz")
    x = torch.tensor(xyz, dtype=torch.float64, requires_grad=True)
    e = scf_energy(x, sp, 5, 5, eps=1e-11, max_iter=200)
    e.backward()
    _, g, _ = scf_gradient(sp, xyz, 5, 5, eps=1e-11, max_iter=200)
    assert np.abs(x.grad.numpy() - g).max() < 1e-12


def test_gradient_refuses_an_unconverged_scf(gpu_lib):
    """max_iter = 3 stops far from the stationary point: the call fails
    (status and message) instead of returning dE/dR of a non-stationary
    density, and the autograd binding raises; the SCF result is still filled."""
    torch = pytest.importorskip("torch")
    import pyfacc
    from pyfacc.grad import scf_energy, scf_gradient
    sp, xyz = load_xyz("h2o_c1.xyz")
    with pytest.raises(pyfacc.PyfaccError, match="did not converge"):
        scf_gradient(sp, xyz, 5, 5, eps=1e-11, max_iter=3)
    x = torch.tensor(xyz, dtype=torch.float32, requires_grad=True)
    with pytest.raises(pyfacc.PyfaccError, match="did not converge"):
        scf_energy(x, sp, 5, 5, eps=1e-11, max_iter=3)
    # converged: the energy is float64 whatever the coordinates' dtype, the
    # gradient comes back in the coordinates' dtype
    e = scf_energy(x, sp, 5, 5, eps=1e-11, max_iter=200)
    assert e.dtype == torch.float64
    e.backward()
    assert x.grad.dtype == torch.float32
    _, g, _ = scf_gradient(sp, xyz, 5, 5, eps=1e-11, max_iter=200)
    assert np.abs(x.grad.double().numpy() - g).max() < 1e-6


@pytest.mark.parametrize("name,na,nb", [("h2o_c1.xyz", 5, 5), ("oh.xyz", 5, 4)])
def test_gradient_terms_vs_oracle_finite_differences(gpu_lib, name, na, nb):
    """Each gradient term against oracle finite differences at fixed D and W:
    E_nuc, Tr D T, Tr D V, -Tr W S and the two-electron energy."""
    import oracle
    from pyfacc.grad import scf_gradient
    sp, xyz = load_xyz(name)
    _, g, r = scf_gradient(sp, xyz, na, nb, eps=1e-11, max_iter=300, want_terms=True)
    D = r["D"]
    uhf = na != nb
    m = oracle.Molecule(sp, xyz)
    S, T, V = m.one_electron()
    Ds = D if uhf else D[0]
    J, K = m.jk_packed(m.eri_packed(), Ds)
    if uhf:
        Jt = J[0] + J[1]
        W = sum(Ds[s] @ (T + V + Jt - K[s]) @ Ds[s] for s in range(2))
        Dt = Ds[0] + Ds[1]
    else:
        W = 0.5 * Ds @ (T + V + J - 0.5 * K) @ Ds
        Dt = Ds

    def parts(coords):
        mm = oracle.Molecule(sp, coords)
        S_, T_, V_ = mm.one_electron()
        J_, K_ = mm.jk_packed(mm.eri_packed(), Ds)
        if uhf:
            e2 = 0.5 * np.sum(Dt * (J_[0] + J_[1])) - 0.5 * sum(np.sum(Ds[s] * K_[s]) for s in range(2))
        else:
            e2 = 0.5 * np.sum(Dt * (J_ - 0.5 * K_))
        return np.array([mm.enuc(), np.sum(Dt * T_), np.sum(Dt * V_), -np.sum(W * S_), e2])

    h = 1e-4
    fd = np.zeros((5, len(sp), 3))
    for a in range(len(sp)):
        for x in range(3):
            cp, cm = xyz.copy(), xyz.copy()
            cp[a, x] += h
            cm[a, x] -= h
            fd[:, a, x] = (parts(cp) - parts(cm)) / (2 * h)
    names = ["nuclear repulsion", "kinetic", "nuclear attraction", "overlap", "two-electron"]
    for k in range(5):
        err = np.abs(r["terms"][k] - fd[k]).max()
        assert err < 1e-6, (names[k], err, r["terms"][k], fd[k])
    assert np.abs(g - fd.sum(axis=0)).max() < 1e-6


# The differentiation workload: the seeded (H2O)32 STO-3G cluster, 224 basis
# functions, 288 gradient components.
def test_gradient_h2o32_vs_finite_differences(gpu_lib):
    from pyfacc.grad import scf_gradient
    sp, xyz = load_xyz("h2o_32.xyz")
    e, g, r = scf_gradient(sp, xyz, 160, 160, eps=1e-11, max_iter=200)
    assert r["converged"] and g.shape == (96, 3)
    assert np.abs(g.sum(axis=0)).max() < 1e-7  # translational invariance over all 288 components
    # a spread sample of components: first, middle and last molecules, O and H
    coords = [(0, 0), (1, 1), (2, 2), (46, 0), (47, 2), (48, 1), (93, 2), (94, 0), (95, 1)]
    fd = _fd(sp, xyz, 160, 160, coords)
    an = np.array([g[a, x] for a, x in coords])
    assert np.abs(an - fd).max() < 1e-6


@pytest.mark.parametrize("na,nb", [(160, 160), (160, 159)])  # RHF, and the (H2O)32+ doublet (UHF)
def test_gradient_h2o32_mixed_forward_matches_fp64(gpu_lib, na, nb):
    from pyfacc.grad import scf_gradient
    sp, xyz = load_xyz("h2o_32.xyz")
    e64, g64, r64 = scf_gradient(sp, xyz, na, nb, eps=1e-10, max_iter=200)
    emx, gmx, rmx = scf_gradient(sp, xyz, na, nb, eps=1e-10, max_iter=200, precision=1)
    assert r64["converged"] and rmx["converged"]
    assert abs(emx - e64) < 1e-6
    assert np.abs(g64.sum(axis=0)).max() < 1e-7
    assert np.abs(g64 - gmx).max() < 1e-6
