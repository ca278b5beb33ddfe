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


@pytest.mark.parametrize("name,na,nb", [("h2o_c1.xyz", 5, 5), ("oh.xyz", 5, 4)])
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
