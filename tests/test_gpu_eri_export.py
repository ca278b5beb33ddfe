"""Shell-quartet ERI blocks in Libint's layout (pf_eri_shell_quartets).

The reference receives one block per shell quartet from Libint
(libint_integrals.f90:111-139: n1 n2 n3 n4 values, first shell slowest,
p components x, y, z) before scattering it into its packed list.  Each GPU
block is compared with the oracle's McMurchie-Davidson block of the same
quartet (a different algorithm), in every orientation, with and without
contraction renormalisation; the blocks are also checked against the packed
export and against the 8-fold permutational symmetry."""
import itertools

import numpy as np
import pytest

from conftest import load_xyz

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,basis,renorm", [("h2o_c1.xyz", "sto-3g", True), ("oh.xyz", "sto-6g", False)])
def test_every_quartet_vs_oracle(gpu_lib, name, basis, renorm):
    import oracle
    import pyfacc
    sp, xyz = load_xyz(name)
    m = oracle.Molecule(sp, xyz, basis, renorm=renorm)
    nsh = m.basis.nsh
    quartets = list(itertools.product(range(nsh), repeat=4))  # all orientations
    blocks = pyfacc.eri_quartets(sp, xyz, quartets, basis, renormalize=renorm)
    worst = 0.0
    for q, blk in zip(quartets, blocks):
        ref = m.shell_eri(*q)
        assert blk.shape == ref.shape, q
        worst = max(worst, np.abs(blk - ref).max())
    assert worst < 1e-12


def test_water_dimer_sample_and_symmetry(gpu_lib):
    import oracle
    import pyfacc
    sp, xyz = load_xyz("h2o_16.xyz")
    sp, xyz = sp[:6], xyz[:6]
    m = oracle.Molecule(sp, xyz)
    nsh = m.basis.nsh
    rng = np.random.default_rng(5)
    quartets = [tuple(int(v) for v in rng.integers(0, nsh, 4)) for _ in range(200)]
    # each with its seven permutational partners
    perms = [lambda a, b, c, d: (b, a, c, d), lambda a, b, c, d: (a, b, d, c), lambda a, b, c, d: (c, d, a, b)]
    full = list(quartets) + [p(*q) for q in quartets for p in perms]
    blocks = pyfacc.eri_quartets(sp, xyz, full)
    n = len(quartets)
    for k, q in enumerate(quartets):
        ref = m.shell_eri(*q)
        assert np.abs(blocks[k] - ref).max() < 1e-12, q
        assert np.abs(blocks[n + 3 * k] - blocks[k].transpose(1, 0, 2, 3)).max() < 1e-14
        assert np.abs(blocks[n + 3 * k + 1] - blocks[k].transpose(0, 1, 3, 2)).max() < 1e-14
        assert np.abs(blocks[n + 3 * k + 2] - blocks[k].transpose(2, 3, 0, 1)).max() < 1e-14


def _packed(i, j, k, l):
    # canonical packed position, the order pf_eri_packed writes (pinned against the oracle elsewhere)
    pr = lambda a, b: a * (a + 1) // 2 + b if a >= b else b * (b + 1) // 2 + a
    ij, kl = pr(i, j), pr(k, l)
    return pr(ij, kl)


def test_blocks_match_packed_export(gpu_lib):
    import pyfacc
    sp, xyz = load_xyz("h2o_c1.xyz")
    packed = pyfacc.eri_packed(sp, xyz)
    info = pyfacc.basis_info(sp, xyz)
    nc = (info["l"] + 1) * (info["l"] + 2) // 2
    first = np.concatenate([[0], np.cumsum(nc)[:-1]])
    nsh = info["nshell"]
    quartets = list(itertools.product(range(nsh), repeat=4))
    blocks = pyfacc.eri_quartets(sp, xyz, quartets)
    for (a, b, c, d), blk in zip(quartets, blocks):
        for i, j, k, l in itertools.product(range(nc[a]), range(nc[b]), range(nc[c]), range(nc[d])):
            idx = _packed(first[a] + i, first[b] + j, first[c] + k, first[d] + l)
            assert abs(blk[i, j, k, l] - packed[idx]) < 1e-13


def test_empty_list_and_bad_index(gpu_lib):
    import pyfacc
    sp, xyz = load_xyz("h2o_c1.xyz")
    assert pyfacc.eri_quartets(sp, xyz, np.zeros((0, 4), dtype=np.int32)) == []
    with pytest.raises(pyfacc.PyfaccError):
        pyfacc.eri_quartets(sp, xyz, [(0, 0, 0, 99)])
