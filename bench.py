#!/usr/bin/env python
"""Benchmark: SCF time-to-converge and Fock builds/s, mixed vs FP64, on one MI355X.

Workload: RHF on the seeded (H2O)64 cluster in STO-3G (448 basis functions,
320 shells; 4x4x4 grid, 3.0 A spacing, random rotations, +-0.2 A jitter,
fixture pyfacc-scf_amd/data/molecules/h2o_64.xyz), direct SCF with Schwarz
threshold 1e-10, DIIS(8), convergence |dE| < 1e-8 and ||dD||_F < 1e-8.

A step is one complete mixed-precision SCF (FP32 incremental J/K on the
difference density with FP64 Fock accumulation, one full FP64 rebuild inside
the FP32 phase, then the switch's full FP64 build and FP32 increments to
convergence under the FP64 stopping rule), started from a
superposition of atomic densities; the same SCF from the reference's
core-Hamiltonian guess is timed beside it.  The same SCF in plain FP64 (full FP64 Fock build every
iteration) is timed the same way for the comparison.  Inputs are built in HBM
by the library before the Fock builds start; every SCF call is synchronous.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pyfacc-scf_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

ANG2BOHR = 1.8897259886
PEAK_FP64_TFLOPS = 78.6
# contracted integrals per shell quartet of each class (Cartesian s = 1, p = 3)
CLASS_ERIS = {"(ss|ss)": 1, "(ps|ss)": 3, "(ps|ps)": 9, "(pp|ss)": 9, "(pp|ps)": 27, "(pp|pp)": 81}


def load_xyz(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0].split()[0])
    sp, xyz = [], []
    for ln in lines[2:2 + n]:
        t = ln.split()
        sp.append(t[0])
        xyz.append([float(t[1]), float(t[2]), float(t[3])])
    return sp, np.array(xyz) * ANG2BOHR


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "jk_direct_h2o64_sto3g_hbm_traffic.md")


def committed_traffic(kernel):
    """HBM bytes per launch of `kernel` (e.g. "jk_direct_kernel (ps|ss) fp32")
    from the committed rocprofv3 --pmc summary: FETCH_SIZE and WRITE_SIZE per
    dispatch, averaged over the kernel's dispatches in the mixed SCF, each
    scaled by what the counter reads for a plain copy of known size.  Byte
    counters cannot be read inside a timed run, so this is the profiled
    figure of the same command, or None if the summary lacks the kernel."""
    try:
        with open(TRAFFIC_SUMMARY) as fh:
            for line in fh:
                if line.startswith("traffic_json:"):
                    table = json.loads(line.split(":", 1)[1])
                    if kernel in table:
                        return float(table[kernel]["bytes_per_launch"]), os.path.relpath(TRAFFIC_SUMMARY, ROOT)
    except OSError:
        pass
    return None, None


def cpu_baseline(species, xyz, n_builds, target_s):
    """Oracle direct Fock build (C, OpenMP over all host threads) on a strided
    sample of the screened shell quartets, scaled to one full build."""
    import oracle
    m = oracle.Molecule(species, xyz)
    threads = oracle.max_threads()
    Q = m.schwarz()
    nocc = int(m.Z.sum()) // 2
    rng = np.random.default_rng(0)
    C = np.linalg.qr(rng.standard_normal((m.nbf, m.nbf)))[0][:, :nocc]
    D = 2.0 * C @ C.T
    def timed(stride, phase=0):
        t0 = time.perf_counter()
        _, _, done = m.jk_direct(D, thresh=1e-10, Q=Q, stride=stride, phase=phase)
        return time.perf_counter() - t0, done

    # loop overhead of visiting every pair of pairs (no integrals), then a
    # calibration sample, then two disjoint samples (every stride-th quartet
    # from offsets 0 and stride/2) of ~target_s/2 seconds each: their mean
    # is the full-build estimate, their difference its sampling spread
    t_loop, _ = timed(1 << 62)
    t_cal, done = timed(512)
    full_est = max(t_cal - t_loop, 1e-3) * 512
    stride = max(2, int(np.ceil(2.0 * full_est / target_s)))
    t_a, done_a = timed(stride, 0)
    t_b, done_b = timed(stride, stride // 2)
    t_cal, done = t_a + t_b, done_a + done_b
    b_a = max(t_a - t_loop, 1e-9) * stride + t_loop
    b_b = max(t_b - t_loop, 1e-9) * stride + t_loop
    t_build = 0.5 * (b_a + b_b)
    # the same build on one thread, on a sparser sample (about target_s / 4);
    # its pair-loop overhead is taken as the all-thread one times the threads
    oracle.set_threads(1)
    try:
        stride1 = max(2, int(np.ceil(t_build * threads / max(target_s / 4, 1.0))))
        t1, done1 = timed(stride1)
    finally:
        oracle.set_threads(threads)
    t_build1 = max(t1 - t_loop * threads, 1e-9) * stride1 + t_loop * threads
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    return {
        "value": 1.0 / (t_build * n_builds),
        "unit": "SCF/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle direct J/K (McMurchie-Davidson, OpenMP {threads} threads) on two disjoint samples, "
                   f"every {stride}-th screened shell quartet from offsets 0 and {stride // 2} ({done} quartets, "
                   f"{t_cal:.1f} s); full build = {t_build:.1f} s (the two samples: {b_a:.1f} s, {b_b:.1f} s), "
                   f"x {n_builds} builds of the FP64 GPU SCF"),
        "fock_build_s": t_build,
        "fock_build_samples_s": [b_a, b_b],
        "sampling_spread": abs(b_a - b_b) / t_build,
        "host": {"nproc": os.cpu_count(), "cpu_model": model},
        "one_thread": {"value": 1.0 / (t_build1 * n_builds), "fock_build_s": t_build1,
                       "sample": f"every {stride1}-th screened shell quartet ({done1} quartets, {t1:.1f} s)",
                       "parallel_speedup": t_build1 / t_build},
    }


def reference_inputs(pyfacc, common):
    """Mixed and FP64 SCFs of three of the reference's own inputs (a
    60-residue glycine chain, 1387 bf; RNA fragments of 4 and 16 nucleotides,
    491 and 1943 bf, as phosphate anions) and of the UHF doublet (H2O)15OH
    (111 bf), one untimed run each first."""
    res = {}
    for name, anion, extra in (("gly60.xyz", False, {}), ("rna1.xyz", True, {}), ("rna4.xyz", True, {}),
                               ("h2o_15oh.xyz", False, {})):
        sp, xyz = load_xyz(os.path.join(ROOT, "pyfacc-scf_amd", "data", "molecules", name))
        # the RNA fragments carry one negative charge per phosphate
        charge = -sp.count("P") if anion else 0
        nel = sum(pyfacc.helper_functions.periodic_table[s] for s in sp) - charge
        nocc, nbeta = (nel + 1) // 2, nel // 2  # the (H2O)15OH doublet runs UHF
        row = {"natoms": len(sp), "charge": charge, "uhf": nocc != nbeta}
        for form, prec in (("mixed", 1), ("fp64", 0)):
            o = pyfacc.default_options(precision=prec, **dict(common, max_iter=150, **extra))
            pyfacc.scf(sp, xyz, nocc, nbeta, options=o)
            t0 = time.perf_counter()
            r = pyfacc.scf(sp, xyz, nocc, nbeta, options=o)
            row[f"scf_time_{form}_s"] = time.perf_counter() - t0
            row[f"iterations_{form}"] = r["iterations"]
            row[f"converged_{form}"] = bool(r["converged"])
            row[f"energy_{form}"] = r["energy"]
            row["nbf"] = r["nbf"]
        row["dE_mixed_vs_fp64"] = row["energy_mixed"] - row["energy_fp64"]
        row["speedup_mixed_vs_fp64"] = row["scf_time_fp64_s"] / row["scf_time_mixed_s"]
        res[name[:-4]] = row
    return res


def gradient_timing(pyfacc, common):
    """SCF + analytic nuclear gradient of (H2O)32 STO-3G (224 bf, 288
    coordinates): wall time of pf_scf_gradient beside pf_scf alone, FP64 and
    with the mixed-precision forward pass (the gradient itself is FP64)."""
    from pyfacc import grad
    sp, xyz = load_xyz(os.path.join(ROOT, "pyfacc-scf_amd", "data", "molecules", "h2o_32.xyz"))
    row = {"natoms": len(sp)}
    for form, prec in (("fp64", 0), ("mixed", 1)):
        o = pyfacc.default_options(precision=prec, **common)
        grad.scf_gradient(sp, xyz, 160, 160, options=o)
        pyfacc.scf(sp, xyz, 160, 160, options=o)
        t_grad = t_scf = 1e30  # best of three of each: the difference of two ~0.05 s walls
        for _ in range(3):
            t0 = time.perf_counter()
            e, g, r = grad.scf_gradient(sp, xyz, 160, 160, options=o)
            t_grad = min(t_grad, time.perf_counter() - t0)
            t0 = time.perf_counter()
            pyfacc.scf(sp, xyz, 160, 160, options=o)
            t_scf = min(t_scf, time.perf_counter() - t0)
        row[f"scf_plus_gradient_s_{form}"] = t_grad
        row[f"scf_s_{form}"] = t_scf
        row[f"gradient_only_s_{form}"] = t_grad - t_scf
        row[f"max_abs_gradient_{form}"] = float(np.abs(g).max())
        row[f"translational_sum_{form}"] = float(np.abs(g.sum(axis=0)).max())
    return row


def batched(pyfacc, count):
    """The batched workload: `count` seeded clusters of 1-12 waters (about 10 %
    with an OH radical, UHF), STO-3G, each to |dE| < 1e-8 and ||dD|| < 1e-8
    from a superposition of atomic densities, in one pf_batch_scf call per
    precision form (setup, stored integrals and all SCF iterations are inside
    the timed call), after one untimed call on the same batch."""
    from pyfacc import clusters
    mols = []
    for i in range(count):
        sp, xyz, _, _ = clusters.batched_instance(i)
        ne = sum(pyfacc.helper_functions.periodic_table[s] for s in sp)
        mols.append((sp, np.asarray(xyz) * ANG2BOHR, (ne + 1) // 2, ne // 2))
    opt = dict(eps=1e-8, max_iter=100, diis=8, schwarz=1e-10, guess=1)  # SAD start, both forms
    # one untimed call on the same batch first: the digest autotune (keyed by
    # batch shape) and other one-time work would otherwise land in whichever
    # form is timed first
    pyfacc.batch_scf(mols, graph=True, precision=0, **opt)
    res = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        t0 = time.perf_counter()
        r = pyfacc.batch_scf(mols, graph=True, precision=prec, **opt)
        res[name] = (time.perf_counter() - t0, r)
    (t64, r64), (tmx, rmx) = res["fp64"], res["mixed"]
    both = r64["converged"] & rmx["converged"]
    closed = np.array([m[2] == m[3] for m in mols])
    st64, stmx = r64["stats"], rmx["stats"]
    return {
        "instances": count,
        "uhf_instances": int(sum(1 for m in mols if m[2] != m[3])),
        "nbf_mean": float(np.mean([5 * m[0].count("O") + m[0].count("H") for m in mols])),  # STO-3G
        "scf_per_s_mixed": count / tmx,
        "scf_per_s_fp64": count / t64,
        "time_mixed_s": tmx,
        "time_fp64_s": t64,
        "scf_loop_mixed_s": stmx["t_scf"],
        "scf_loop_fp64_s": st64["t_scf"],
        "setup_s": st64["t_setup"],
        "eri_build_s": st64["t_eri"],
        "stored_integrals": st64["eri_values"],
        "store_gb": st64["store_bytes"] / 1e9,
        "iterations_fp64": int(st64["iterations"]),
        "iterations_mixed": int(stmx["iterations"]),
        "converged_fp64": int(r64["converged"].sum()),
        "converged_mixed": int(rmx["converged"].sum()),
        # over the molecules converged in both forms (the others stop at max_iter)
        "max_abs_dE_mixed_vs_fp64": float(np.abs(rmx["energy"] - r64["energy"])[both].max()),
        # instances whose two precisions land more than 1e-6 Eh apart: UHF ones
        # settling on different stationary points (not rounding), split by kind
        "n_abs_dE_above_1e-6": int((np.abs(rmx["energy"] - r64["energy"])[both] > 1e-6).sum()),
        "max_abs_dE_mixed_vs_fp64_rhf": float(np.abs(rmx["energy"] - r64["energy"])[both & closed].max()),
        "n_abs_dE_above_1e-6_uhf": int((np.abs(rmx["energy"] - r64["energy"])[both & ~closed] > 1e-6).sum()),
        "guess": "SAD",
        "hipgraph": bool(st64["graph"]),
        "digest_waves_autotuned": int(st64["digest_waves"]),
        "digest_item_values_autotuned": float(st64["digest_item_values"]),
        "digest_tail_waves_autotuned": int(st64["digest_tail_waves"]),
        "digest_tail_item_values_autotuned": float(st64["digest_tail_item_values"]),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--xyz", default=os.path.join(ROOT, "pyfacc-scf_amd", "data", "molecules", "h2o_64.xyz"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--batch", type=int, default=4096,
                    help="instances of the batched water-cluster workload (0: skip)")
    ap.add_argument("--no-reference-inputs", action="store_true",
                    help="skip the SCFs of the reference's glycine-chain and RNA inputs")
    ap.add_argument("--jk-only", action="store_true",
                    help="only one FP64 and one FP32 full direct J/K build of the workload at a seeded RHF "
                         "density (after one untimed build each): the short command for rocprofv3 counter runs")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (the byte-counter "
                         "calibration kernel for rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "design is described but not run", file=sys.stderr)
        sys.exit(2)

    import pyfacc
    from pyfacc import flops

    if pyfacc.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_mib > 0:
        pyfacc._native.device_copy(args.copy_mib << 20)
    species, xyz = load_xyz(args.xyz)
    nelec = sum(pyfacc.helper_functions.periodic_table[s] for s in species)
    nocc = nelec // 2
    if args.jk_only:
        nbf = pyfacc.basis_info(species, xyz)["nbf"]
        rng = np.random.default_rng(7)
        C = np.linalg.qr(rng.standard_normal((nbf, nbf)))[0][:, :nocc]
        D = 2.0 * C @ C.T
        out = {}
        for fp32 in (False, True):
            pyfacc.jk(species, xyz, D, schwarz=1e-10, fp32=fp32)
            t0 = time.perf_counter()
            J, K = pyfacc.jk(species, xyz, D, schwarz=1e-10, fp32=fp32)
            out["fp32" if fp32 else "fp64"] = {"s": time.perf_counter() - t0, "J_fro": float(np.linalg.norm(J))}
        print(json.dumps({"jk_only": out, "nbf": nbf}))
        return
    # superposition-of-atomic-densities start for both forms (the core guess of
    # the reference is timed beside them below)
    common = dict(eps=1e-8, max_iter=100, diis=8, schwarz=1e-10, guess=1)
    opt64 = pyfacc.default_options(precision=0, **common)
    optmx = pyfacc.default_options(precision=1, **common)
    optinc = pyfacc.default_options(precision=0, incremental=1, **common)

    for _ in range(args.warmup):
        pyfacc.scf(species, xyz, nocc, nocc, options=optmx)
        pyfacc.scf(species, xyz, nocc, nocc, options=opt64)

    # ---- timed region: K mixed-precision SCF solutions (no per-class timing
    # events inside: pf_options.profile = 0)
    t0 = time.perf_counter()
    mixed = [pyfacc.scf(species, xyz, nocc, nocc, options=optmx) for _ in range(args.steps)]
    t_mixed = time.perf_counter() - t0

    # ---- comparison: plain FP64 and FP64-incremental, timed the same way
    t0 = time.perf_counter()
    fp64 = [pyfacc.scf(species, xyz, nocc, nocc, options=opt64) for _ in range(args.steps)]
    t_fp64 = time.perf_counter() - t0

    # ---- the same two SCFs once more with HIP events around every J/K class
    # kernel (on the engine's stream): the per-class kernel times of the
    # roofline figure, kept out of the timed SCFs above
    prof_mx = pyfacc.scf(species, xyz, nocc, nocc, options=pyfacc.default_options(precision=1, profile=1, **common))
    q_mx, ms_mx, n_mx, orient, p_mx = pyfacc.scf_profile(with_prims=True)
    prof_64 = pyfacc.scf(species, xyz, nocc, nocc, options=pyfacc.default_options(precision=0, profile=1, **common))
    q_64, ms_64, n_64, orient64, p_64 = pyfacc.scf_profile(with_prims=True)
    t0 = time.perf_counter()
    inc = pyfacc.scf(species, xyz, nocc, nocc, options=optinc)
    t_inc = time.perf_counter() - t0
    # the same two forms from the core-Hamiltonian guess (rscf.f90:113-114)
    core = {}
    for name, prec in (("mixed", 1), ("fp64", 0)):
        o = pyfacc.default_options(precision=prec, **dict(common, guess=0))
        t0 = time.perf_counter()
        rc = pyfacc.scf(species, xyz, nocc, nocc, options=o)
        core[name] = (time.perf_counter() - t0, rc)

    e64 = fp64[-1]["energy"]
    emx = mixed[-1]["energy"]
    # the CPU reference energy of the headline cluster: the oracle's energy
    # functional at a stationary density of its own SCF equations (one oracle
    # Fock build; tests/golden/make_h2o64_certificate.py)
    e_cpu = None
    cert = os.path.join(ROOT, "tests", "golden", "h2o64_certificate.json")
    if os.path.abspath(args.xyz) == os.path.join(ROOT, "pyfacc-scf_amd", "data", "molecules", "h2o_64.xyz") and \
            os.path.exists(cert):
        with open(cert) as fh:
            e_cpu = json.load(fh)["energy_functional"]
    # ---- roofline of the dominant kernel of the timed region (mixed SCF)
    flat_ms = ms_mx.reshape(-1)
    k = int(np.argmax(flat_ms))
    prec, cls = divmod(k, 6)
    launches = int(n_mx[prec, cls])
    avg_ms = flat_ms[k] / max(launches, 1)
    q_per_launch = q_mx[prec, cls] / max(launches, 1)
    fl = flops.class_flops(cls, q_per_launch, ndens=1, taylor_terms=7 if prec == 0 else 5, orient=orient[cls],
                           prim_quartets=p_mx[prec, cls] / max(launches, 1))
    achieved = fl / (avg_ms * 1e-3) / 1e12
    kname = f"jk_direct_kernel {flops.CLASS_NAMES[cls]} {'fp64' if prec == 0 else 'fp32'}"
    traffic, traffic_src = committed_traffic(kname)
    # the same kernel class in the plain FP64 SCF (full builds)
    k64 = int(np.argmax(ms_64[0]))
    avg64 = ms_64[0, k64] / max(n_64[0, k64], 1)
    nb64 = np.maximum(n_64[0], 1)
    fl64 = flops.class_flops(k64, q_64[0, k64] / nb64[k64], orient=orient64[k64],
                             prim_quartets=p_64[0, k64] / nb64[k64])
    fock64_flops = flops.fock_flops(q_64[0] / nb64, orients=orient64, class_prims=p_64[0] / nb64)
    per_class = {flops.CLASS_NAMES[c]: {
        "orient": "".join("sp"[int(x)] for x in orient64[c]),
        "quartets_per_build": int(q_64[0, c] / max(n_64[0, c], 1)),
        "ms_per_build": float(ms_64[0, c] / max(n_64[0, c], 1)),
        "prim_quartets_per_build": int(p_64[0, c] / nb64[c]),
        "tflops": float(flops.class_flops(c, q_64[0, c] / nb64[c], orient=orient64[c],
                                          prim_quartets=p_64[0, c] / nb64[c]) /
                        max(ms_64[0, c] / nb64[c] * 1e-3, 1e-12) / 1e12)} for c in range(6)}
    for v in per_class.values():
        v["frac"] = v["tflops"] / PEAK_FP64_TFLOPS
    # every J/K class of the timed mixed SCF (FP32 increments and the FP64
    # anchor and switch builds), from the profiled mixed SCF
    mixed_per_class = {}
    for pr in (0, 1):
        for c in range(6):
            nl = int(n_mx[pr, c])
            if nl == 0:
                continue
            ms = float(ms_mx[pr, c] / nl)
            f = flops.class_flops(c, q_mx[pr, c] / nl, ndens=1, taylor_terms=7 if pr == 0 else 5, orient=orient[c],
                                  prim_quartets=p_mx[pr, c] / nl)
            tf = f / (ms * 1e-3) / 1e12
            mixed_per_class[f"{flops.CLASS_NAMES[c]} {'fp64' if pr == 0 else 'fp32'}"] = {
                "launches": nl, "avg_launch_ms": ms, "tflops": tf,
                "frac": tf / (PEAK_FP64_TFLOPS if pr == 0 else 157.3),
                "frac_of_unpacked_fp32_valu": tf / PEAK_FP64_TFLOPS if pr == 1 else None}

    out = {
        "metric": "SCF time-to-converge & Fock builds/s, mixed vs FP64 GPU at |dE|<=1e-6 Eh",
        "value": args.steps / t_mixed,
        "unit": "SCF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64 (mixed)",
        "data": "synthetic: seeded (H2O)64 cluster geometry, STO-3G (BSE data)",
        "config": {
            "workload": "(H2O)64 STO-3G RHF direct SCF, 448 bf, Schwarz 1e-10, DIIS 8, eps 1e-8, SAD guess",
            "nbf": int(mixed[-1]["nbf"]), "nshell": int(mixed[-1]["nshell"]), "parallelism": "single GPU",
        },
        "scf_time_mixed_s": t_mixed / args.steps,
        "scf_time_fp64_s": t_fp64 / args.steps,
        "scf_time_fp64_incremental_s": t_inc,
        "core_guess": {"scf_time_mixed_s": core["mixed"][0], "scf_time_fp64_s": core["fp64"][0],
                       "iterations_mixed": core["mixed"][1]["iterations"],
                       "iterations_fp64": core["fp64"][1]["iterations"],
                       "dE_mixed_vs_fp64": core["mixed"][1]["energy"] - core["fp64"][1]["energy"]},
        "speedup_mixed_vs_fp64": (t_fp64 / args.steps) / (t_mixed / args.steps),
        "dE_mixed_vs_fp64": emx - e64,
        "dE_mixed_vs_cpu_reference": None if e_cpu is None else emx - e_cpu,
        "dE_fp64_vs_cpu_reference": None if e_cpu is None else e64 - e_cpu,
        "energy_fp64": e64,
        "iterations_fp64": fp64[-1]["iterations"],
        "iterations_mixed": mixed[-1]["iterations"],
        "fock_builds_fp64_path": fp64[-1]["fock_builds_fp64"],
        "fock_builds_mixed_path": [mixed[-1]["fock_builds_fp32"], mixed[-1]["fock_builds_fp64"]],
        # build times from the profiled FP64 SCF (per-iteration events only when profiling)
        "fock_builds_per_s_fp64": prof_64["fock_builds_fp64"] / prof_64["t_fock"],
        "fock_build_ms_fp64": 1e3 * prof_64["t_fock"] / prof_64["fock_builds_fp64"],
        "fock_build_tflops_fp64": fock64_flops / (prof_64["t_fock"] / prof_64["fock_builds_fp64"]) / 1e12,
        "converged": bool(mixed[-1]["converged"] and fp64[-1]["converged"]),
        # throughput of one FP64 full Fock build (J/K kernel time of its six classes)
        "fp64_build_throughput": {
            "shell_quartets_per_s": float(sum(v["quartets_per_build"] for v in per_class.values()) /
                                          (sum(v["ms_per_build"] for v in per_class.values()) * 1e-3)),
            "contracted_eris_per_s": float(sum(v["quartets_per_build"] * CLASS_ERIS[c] for c, v in per_class.items()) /
                                           (sum(v["ms_per_build"] for v in per_class.values()) * 1e-3)),
            "primitive_quartets_per_s": float(sum(v["prim_quartets_per_build"] for v in per_class.values()) /
                                              (sum(v["ms_per_build"] for v in per_class.values()) * 1e-3)),
        },
        "fp64_build_per_class": per_class,
        "mixed_scf_per_class": mixed_per_class,
        "roofline": {
            # the J/K kernels are vector-ALU (VALU) FP64/FP32 arithmetic: no MFMA
            "bound": "valu",
            "kernel": kname,
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS if prec == 0 else 157.3,
            "unit": "TFLOP/s",
            "frac": achieved / (PEAK_FP64_TFLOPS if prec == 0 else 157.3),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "avg_launch_ms": avg_ms,
            "launches": launches,
            "timing_source": "HIP events around each J/K class kernel of one profiled mixed SCF "
                             "(pf_options.profile = 1), run after the timed SCFs",
            "profiled_scf_s": {"mixed": prof_mx["t_total"], "fp64": prof_64["t_total"]},
            "flops_per_launch": fl,
            # FP32 kernels: the build compiles with packed FP32 disabled, so the
            # FP32 VALU rate it can reach is the unpacked one, equal to FP64's
            "frac_of_unpacked_fp32_valu": achieved / PEAK_FP64_TFLOPS if prec == 1 else None,
            "fp64_full_build_dominant": {"kernel": flops.CLASS_NAMES[k64], "avg_launch_ms": avg64,
                                         "tflops": fl64 / (avg64 * 1e-3) / 1e12,
                                         "frac": fl64 / (avg64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS},
        },
    }
    # the metric's accuracy bar: both SCFs converged and |E_mixed - E_FP64| <= 1e-6 Eh
    bar_met = bool(all(r["converged"] for r in mixed + fp64) and
                   max(abs(r["energy"] - e64) for r in mixed) <= 1e-6)
    out["bar_met"] = bar_met
    if not bar_met:
        out["value"] = None
    if not args.no_reference_inputs:
        out["reference_inputs"] = reference_inputs(pyfacc, common)
        out["gradient_h2o32"] = gradient_timing(pyfacc, common)
    if args.batch > 0:
        out["batched"] = batched(pyfacc, args.batch)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(species, xyz, fp64[-1]["fock_builds_fp64"], args.cpu_seconds)
    print(json.dumps(out))
    if not bar_met:
        print("bench.py: the mixed SCF missed the |dE| <= 1e-6 Eh bar or did not converge", file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
