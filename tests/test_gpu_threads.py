"""Concurrent pf_scf calls from several host threads.  ctypes releases the
GIL around the call, so the SCFs overlap on the device, each on its own
stream with its own buffers; the caches they share (parsed basis files,
atomic densities of the SAD start, the J'_cd scratch pool) are mutex-guarded.
Every thread's energy must equal the same SCF run alone (within 1e-8 Eh:
FP64 atomics sum in no fixed order, so repeated runs differ at ~1e-11)."""
import threading

import pytest

from conftest import load_xyz

pytestmark = pytest.mark.gpu


def test_concurrent_scf_calls_match_sequential(gpu_lib):
    import pyfacc
    sp16, xyz16 = load_xyz("h2o_16.xyz")
    sp15, xyz15 = load_xyz("h2o_15oh.xyz")
    jobs = [
        (sp16, xyz16, 80, 80, {"precision": 0}),
        (sp16, xyz16, 80, 80, {"precision": 1}),
        (sp15, xyz15, 80, 79, {"precision": 0}),
        (sp16[:12], xyz16[:12], 20, 20, {"basis": "sto-6g"}),
    ]
    kw = dict(eps=1e-9, max_iter=150, diis=8, guess=1)
    ref = []
    for sp, xyz, na, nb, o in jobs:
        r = pyfacc.scf(sp, xyz, na, nb, **kw, **o)
        assert r["converged"]
        ref.append(r["energy"])
    tasks = [i for i in range(len(jobs)) for _ in range(2)]  # two copies of each at once
    out = [None] * len(tasks)
    errors = []

    def run(t):
        sp, xyz, na, nb, o = jobs[tasks[t]]
        try:
            out[t] = pyfacc.scf(sp, xyz, na, nb, **kw, **o)
        except Exception as e:  # reported below, with the task that raised
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=run, args=(t,)) for t in range(len(tasks))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "an SCF thread did not finish"
    assert not errors, errors
    for t, r in enumerate(out):
        assert r["converged"], t
        assert abs(r["energy"] - ref[tasks[t]]) < 1e-8, (t, r["energy"], ref[tasks[t]])


def test_short_lived_threads_reuse_cached_device_memory(gpu_lib):
    """Threads that run one SCF each and end hand their cached device blocks
    to a shared idle list that later threads reuse: the cache stays the size
    of one SCF's working set however many threads come and go, and
    release_device_cache empties it."""
    import pyfacc
    sp, xyz = load_xyz("h2o_16.xyz")
    kw = dict(eps=1e-8, max_iter=150, diis=8, guess=1)
    e0 = pyfacc.scf(sp, xyz, 80, 80, **kw)["energy"]
    sizes = []
    for k in range(6):
        out = {}
        th = threading.Thread(target=lambda: out.setdefault("r", pyfacc.scf(sp, xyz, 80, 80, **kw)))
        th.start()
        th.join(timeout=100)
        assert not th.is_alive()
        assert abs(out["r"]["energy"] - e0) < 1e-8
        sizes.append(pyfacc.cached_device_bytes())
    assert sizes[1] > 0
    assert max(sizes[2:]) <= sizes[1], sizes  # no growth after the first thread's blocks were handed over
    pyfacc.release_device_cache()
    assert pyfacc.cached_device_bytes() == 0
    assert abs(pyfacc.scf(sp, xyz, 80, 80, **kw)["energy"] - e0) < 1e-8  # the library still works after it
