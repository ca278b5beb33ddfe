"""Oracle certificate of the (H2O)64 STO-3G RHF solution (448 basis functions).

A full oracle SCF of the headline cluster takes hours on a few host cores,
one oracle Fock build about a quarter of an hour.  So the oracle checks a
candidate instead of solving: given a converged RHF density D (an .npy file,
nbf x nbf, e.g. written by the GPU path at the bench settings), it builds
F = H + J[D] - K[D]/2 with its own integrals (McMurchie-Davidson, Schwarz
1e-12 on Q_ab Q_cd) and reports

* the stationarity residual max |F D S - S D F| (zero exactly at an SCF
  solution of the oracle's equations),
* Tr(D S) and max |D S D - 2 D| (an RHF density of 640 electrons),
* the oracle's energy functional E[D] = 1/2 sum D (H + F) + E_nuc.

At a density within delta of the oracle's solution, E[D] differs from the
oracle's SCF energy by O(delta^2), so E[D] is the CPU reference energy of
the cluster to far below 1e-9 Eh when the residual is ~1e-8.  Only the
figures are committed (h2o64_certificate.json), not the density.

Usage: python tests/golden/make_h2o64_certificate.py D.npy
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from conftest import load_xyz  # noqa: E402


def main(path):
    D = np.load(path)
    sp, xyz = load_xyz("h2o_64.xyz")
    m = oracle.Molecule(sp, xyz)
    assert D.shape == (m.nbf, m.nbf)
    S, T, V = m.one_electron()
    H = T + V
    t0 = time.perf_counter()
    J, K, nq = m.jk_direct(D, thresh=1e-12)
    t_build = time.perf_counter() - t0
    F = H + J - 0.5 * K
    comm = F @ D @ S - S @ D @ F
    out = {
        "system": "(H2O)64 STO-3G RHF, h2o_64.xyz, 448 bf, renormalised contractions",
        "oracle_schwarz": 1e-12,
        "oracle_quartets": int(nq),
        "oracle_build_s": t_build,
        "oracle_threads": oracle.max_threads(),
        "energy_functional": float(0.5 * np.sum(D * (H + F)) + m.enuc()),
        "e_nuc": float(m.enuc()),
        "stationarity_residual_max": float(np.abs(comm).max()),
        "trace_DS": float(np.trace(D @ S)),
        "idempotency_max": float(np.abs(D @ S @ D - 2.0 * D).max()),
        "density_source": os.path.basename(path),
    }
    with open(os.path.join(HERE, "h2o64_certificate.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1])
