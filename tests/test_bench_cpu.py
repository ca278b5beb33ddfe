"""bench.py's CPU-baseline leg (the oracle's direct Fock build on a strided
sample of quartets, all threads and one thread) runs on the host alone;
checked here on the water monomer with a short sample."""
import os
import sys

from conftest import ROOT, load_xyz


def test_cpu_baseline_leg_reports_contract_fields():
    sys.path.insert(0, ROOT)
    import bench
    sp, xyz = load_xyz("h2o_c1.xyz")
    cb = bench.cpu_baseline(sp, xyz, n_builds=10, target_s=0.2)
    assert cb["unit"] == "SCF/s" and cb["kind"] == "port"
    assert cb["value"] > 0 and cb["fock_build_s"] > 0 and cb["cores"] >= 1
    assert abs(cb["value"] * cb["fock_build_s"] * 10 - 1.0) < 1e-9
    assert cb["host"]["nproc"] == os.cpu_count()
    assert cb["one_thread"]["fock_build_s"] > 0
