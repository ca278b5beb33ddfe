"""The full 4096-instance batched workload (1-12 waters, about 10 % with an
OH radical, SAD start, Schwarz 1e-10, eps 1e-8) in both precisions: every
instance converges, and every instance, RHF and UHF alike, ends within the
1e-6 Eh bar of its FP64 energy (the FP64 opening phase of the mixed form
keeps the open-shell instances on the FP64 solution)."""
import numpy as np
import pytest

from conftest import ANG2BOHR

pytestmark = pytest.mark.gpu


def test_batch_4096_mixed_vs_fp64(gpu_lib):
    import pyfacc
    from pyfacc import clusters
    mols = []
    for i in range(4096):
        sp, xyz, _, _ = clusters.batched_instance(i)
        ne = sum({"H": 1, "O": 8}[s] for s in sp)
        mols.append((sp, np.asarray(xyz) * ANG2BOHR, (ne + 1) // 2, ne // 2))
    uhf = np.array([na != nb for _, _, na, nb in mols])
    assert 300 < uhf.sum() < 500
    opt = dict(eps=1e-8, max_iter=100, diis=8, schwarz=1e-10, guess=1)
    r64 = pyfacc.batch_scf(mols, graph=True, precision=0, **opt)
    mx = pyfacc.batch_scf(mols, graph=True, precision=1, **opt)
    assert r64["converged"].all() and mx["converged"].all()
    d = np.abs(mx["energy"] - r64["energy"])
    assert d[~uhf].max() < 1e-6
    assert d[uhf].max() < 1e-6, [int(i) for i in np.where(d > 1e-6)[0]]
    assert mx["stats"]["graph"] == 1 and r64["stats"]["graph"] == 1
