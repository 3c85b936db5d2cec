"""Benchmark of the SQ density step (the spectral-quadrature hot path).

A step is one ClenshawCurtisSpectralQuadrature call (sq.cpp:377-834) over
every FD node of the workload: Lanczos bounds, Chebyshev moments, Fermi
level, rho/Ebs/Eent.  Inputs (Veff, projector tables) are resident on the
device before the timed region; the call returns after its stream has
synchronised.  Prints one JSON line.

Default workload: rocksalt LiH 3x3x3 (216 atoms), 48^3 FD nodes, 12th-order
FD, T = 40,000 K, Rcut = 7 bohr (nloc = 15, 31^3 cube), npl = 64, FP64.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SQ nodes/sec (mixed vs FP64 at rel rho err <=1e-6); SCF step s; % HBM peak"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "lih216": ("rocksalt LiH 3x3x3 (216 atoms), 48^3 nodes, FD order 12, nloc=15 (31^3 cube), "
               "npl=64, T=40000 K, seeded 0.05 bohr displacements"),
    "al32": ("FCC Al 2x2x2 (32 atoms), 32^3 nodes, FD order 12, nloc=17 (35^3 cube), npl=48, "
             "T=58000 K, seeded 0.1 bohr displacements"),
    "al256": ("FCC Al 4x4x4 (256 atoms), 64^3 nodes, FD order 12, nloc=17 (35^3 cube), npl=48, "
              "T=58000 K, seeded 0.1 bohr displacements"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return "unknown"


def host_threads() -> int:
    """The CPU share this process may use: its affinity set, capped by
    OMP_NUM_THREADS when that is set (16 per GPU on the GPU hosts)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(n, 1)


def cpu_baseline(system, veff, Rcut, T, seconds, seed=5):
    """Times the FP64 oracle's per-node loop (an -O3 build of the same source,
    OpenMP over nodes on every thread of this process's CPU share) on a seeded
    node sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # noqa: E402
    import helpers  # noqa: E402

    libpath = oracle.use_timing_build()
    osys = helpers.oracle_for(system, Rcut, T)
    threads = host_threads()
    rng = np.random.Generator(np.random.PCG64(seed))
    # calibrate on one node per thread, then size the sample to ~`seconds`
    probe = rng.choice(system.nn, size=threads, replace=False).astype(np.int32)
    t0 = time.perf_counter()
    osys.moments_nodes(veff, probe, nthreads=threads)
    tp = time.perf_counter() - t0
    per_round = max(tp, 1e-3)
    rounds = max(1, int(seconds / per_round))
    sample = rng.choice(system.nn, size=min(system.nn, rounds * threads), replace=False).astype(np.int32)
    t0 = time.perf_counter()
    osys.moments_nodes(veff, sample, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": sample.size / dt, "unit": "nodes/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "build": os.path.basename(libpath) + " (gcc -O3 -funroll-loops, "
            + ("-march=x86-64-v4" if libpath.endswith("_v4.so") else "-march=x86-64-v3") + ", OpenMP)",
            "sample": f"{sample.size} seeded random nodes of the same workload (Lanczos + Chebyshev moments "
                      f"per node, FP64 oracle sq_oracle.c, OpenMP over nodes on {threads} threads), {dt:.1f} s; "
                      f"the Fermi solve and rho assembly (<1% of the step) are excluded",
            "seconds": dt}


def profiled_traffic(workload: str, precision: str) -> dict:
    """HBM bytes per dispatch of the z-march kernels as measured by the counter
    runs recorded under profiles/ (lines `traffic <workload> <precision>
    <kernel> <bytes>`); empty when no run covers this workload."""
    out = {}
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return out
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".md"):
            continue
        with open(os.path.join(pdir, fn)) as f:
            for line in f:
                t = line.split()
                if len(t) == 5 and t[0] == "traffic" and t[1] == workload and t[2] == precision:
                    out[t[3]] = float(t[4])
    return out


def calibration_copy(nb: int) -> None:
    """One device-to-device copy of nb bytes through the HIP runtime (the
    runtime's copy kernel): nb bytes read and nb written, the reference that
    scales the FETCH_SIZE / WRITE_SIZE counters of the other kernels."""
    import ctypes as C
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    a, b = C.c_void_p(), C.c_void_p()
    for call in (lambda: hip.hipMalloc(C.byref(a), C.c_size_t(nb)), lambda: hip.hipMalloc(C.byref(b), C.c_size_t(nb)),
                 lambda: hip.hipMemset(a, 1, C.c_size_t(nb)), lambda: hip.hipDeviceSynchronize(),
                 lambda: hip.hipMemcpy(b, a, C.c_size_t(nb), 3), lambda: hip.hipDeviceSynchronize(),
                 lambda: hip.hipFree(a), lambda: hip.hipFree(b)):
        if call() != 0:
            raise RuntimeError("calibration copy: a HIP runtime call failed")
    log(f"bench: calibration copy of {nb} bytes read + {nb} bytes written")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lih216", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--batch-nodes", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="before the steps, run one plain device copy of this size (MiB) so that "
                         "counter runs can scale FETCH_SIZE/WRITE_SIZE by a kernel of known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); "
              "the SQ node loop shards by node bricks with one Veff all-gather per SCF step, "
              "which is not run here", file=sys.stderr)
        sys.exit(2)

    from sqdft_amd import configs
    from sqdft_amd.sq import SpectralQuadrature

    cfg = configs.CONFIGS[args.workload]
    log(f"bench: building {args.workload}")
    S = configs.build(args.workload, seed=0)
    veff = configs.smooth_veff(S)
    sq = SpectralQuadrature(S, precision=args.precision, batch_nodes=args.batch_nodes)
    info = S.info
    nn = S.nn
    if args.copy_calib_mib > 0:
        calibration_copy(args.copy_calib_mib * (1 << 20))
    for i in range(args.warmup):
        log(f"bench: warmup {i}")
        sq.density(veff, scf_iter=1)
    times, last = [], None
    for i in range(args.steps):
        t0 = time.perf_counter()
        last = sq.density(veff, scf_iter=1)
        times.append(time.perf_counter() - t0)
        log(f"bench: step {i} {times[-1]:.3f} s lambda_f={last['lambda_f']:.12f}")
    ms = 1e3 * float(np.mean(times))
    value = nn / (ms * 1e-3)
    st = last["stats"]
    # dominant z-march kernel: HIP events around every launch of the last step
    # hot kernels, timed with HIP events around every launch of the last step
    kt = [sq.kernel_timing(w) for w in range(4)]
    keys = ["k_zmarch<CHEB>", "k_zmarch<LANCZOS>", "k_projs<LANCZOS>", "k_projs<CHEB>"]
    what = ["Chebyshev recurrence: read t_k, t_{k-1}, write t_{k+1}",
            "Lanczos pass: read r', v_{k-1}, write v_k, r'_{k+1}",
            "projections of the Lanczos vector: read r', v_{k-1}",
            "projections of t_k: read t_k"]
    nvec = [3, 4, 2, 1]
    dom = int(np.argmax([k["total_ms"] for k in kt]))
    rs = 8 if args.precision == "fp64" else 4
    N = (2 * info["nloc"] + 1) ** 3
    traffic = profiled_traffic(args.workload, args.precision)
    rl = {}
    for w in range(4):
        k = kt[w]
        avg_ms = k["total_ms"] / max(k["launches"], 1)
        alg = k["alg_bytes"] / max(k["launches"], 1)
        rl[w] = {"kernel": f"{keys[w]} ({what[w]})", "launches": k["launches"], "avg_ms": avg_ms,
                 "alg_bytes_per_launch": alg, "achieved_GBs": alg / (avg_ms * 1e-3) / 1e9 if avg_ms else 0.0,
                 "alg_bytes_per_unit": f"{nvec[w]} * {N} * {rs} B per node-step",
                 "traffic": traffic.get(keys[w])}
    roof = {"bound": "hbm", "kernel": rl[dom]["kernel"], "achieved": rl[dom]["achieved_GBs"],
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rl[dom]["achieved_GBs"] / HBM_PEAK_GBS,
            "traffic": rl[dom]["traffic"],
            "alg_bytes_per_unit": rl[dom]["alg_bytes_per_unit"],
            "traffic_source": "rocprofv3 FETCH_SIZE / WRITE_SIZE per dispatch, scaled by a plain copy "
                              "(profiles/traffic_lih216.md)" if rl[dom]["traffic"] else None,
            "others": [rl[w] for w in range(4) if w != dom]}
    out = {"metric": METRIC, "value": value, "unit": "nodes/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
           "data": "synthetic (seeded atomic positions; smooth model Veff)",
           "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}", "nodes": nn,
                      "nloc": info["nloc"], "npl": info["npl"], "precision": args.precision,
                      "batch_nodes": args.batch_nodes or nn},
           "phases_s": {"lanczos": st["t_lanczos"], "cheb": st["t_cheb"], "fermi": st["t_fermi"],
                        "rho": st["t_rho"]},
           "lanczos_iters": {"node0": st["lanczos_iters_node0"], "min": st["lanczos_iters_min"],
                             "max": st["lanczos_iters_max"], "passes": st["lanczos_passes"]},
           "roofline": roof}
    if not args.no_mixed and args.precision == "fp64":
        log("bench: mixed precision step")
        sqm = SpectralQuadrature(S, precision="mixed", batch_nodes=args.batch_nodes)
        sqm.density(veff, scf_iter=1)
        t0 = time.perf_counter()
        mo = sqm.density(veff, scf_iter=1)
        tm = time.perf_counter() - t0
        err = float(np.max(np.abs(mo["rho"] - last["rho"])) / np.max(np.abs(last["rho"])))
        sqm.close()
        # accuracy of the mixed arithmetic itself: both Fermi levels solved to
        # 1e-12 (at fermi_tol = 1e-6 the two roots may differ by ~1e-6, which
        # alone moves rho by a few 1e-6)
        log("bench: mixed vs fp64 accuracy at fermi_tol 1e-12")
        r64 = SpectralQuadrature(S, precision="fp64", batch_nodes=args.batch_nodes, fermi_tol=1e-12)
        t64 = r64.density(veff, scf_iter=1)
        f64 = r64.ForcesNonLocal()
        r64.close()
        r32 = SpectralQuadrature(S, precision="mixed", batch_nodes=args.batch_nodes, fermi_tol=1e-12)
        t32 = r32.density(veff, scf_iter=1)
        f32 = r32.ForcesNonLocal()
        r32.close()
        err_t = float(np.max(np.abs(t32["rho"] - t64["rho"])) / np.max(np.abs(t64["rho"])))
        out["mixed"] = {"value": nn / tm, "unit": "nodes/s", "ms_per_step": 1e3 * tm,
                        "speedup_vs_fp64": (ms * 1e-3) / tm,
                        "rho_rel_err_vs_fp64": err_t,
                        "lambda_f_err": abs(t32["lambda_f"] - t64["lambda_f"]),
                        "rho_rel_err_at_fermi_tol_1e-6": err,
                        "ebs_rel_err": abs(t32["ebs"] - t64["ebs"]) / abs(t64["ebs"]),
                        "nonlocal_force_max_err": float(np.max(np.abs(f32 - f64))),
                        "nonlocal_force_max": float(np.max(np.abs(f64)))}
    sq.close()
    if not args.no_scf:
        # one SCF iteration on the workload's own pseudocharge: SQ density,
        # total energy, AAJ Poisson from the previous phi, LDA potential and the
        # Veff residual (src/scf.cpp:145-200); the second iteration is timed
        log("bench: scf iterations")
        from sqdft_amd.scf import SCF
        scf = SCF(S, precision=args.precision, batch_nodes=args.batch_nodes)
        t0 = time.perf_counter()
        scf.Pseudocharge()
        t_pc = time.perf_counter() - t0
        scf.initial_potential()
        it1 = scf.iteration()
        scf.veff = it1["veff_new"]
        it2 = scf.iteration()
        log("bench: forces")
        t0 = time.perf_counter()
        scf.ForcesLocal()
        t_loc = time.perf_counter() - t0
        t0 = time.perf_counter()
        scf.sq.ForcesNonLocal()
        t_nl = time.perf_counter() - t0
        scf.close()
        out["scf_step"] = {"s": it2["t_sq"] + it2["t_rest"], "sq_s": it2["t_sq"],
                           "energy_poisson_vxc_s": it2["t_rest"], "poisson_iters": it2["poisson"][0],
                           "scf_err": it2["scf_err"], "lambda_f": it2["lambda_f"],
                           "forces_local_s": t_loc, "forces_nonlocal_s": t_nl, "pseudocharge_s": t_pc,
                           "what": "second SCF iteration from the atomic starting density, unmixed Veff; "
                                   "then local and nonlocal forces at that state"}
    if not args.no_scf:
        # SCF iterations to SCF_err 1e-6 on the reference's Al test input:
        # its Periodic Pulay mixing vs Newton steps from the adjoint response
        log("bench: scf mixing on the Al test input")
        from sqdft_amd.scf import SCF
        from sqdft_amd.system import System
        A = System.from_files(os.path.join(ROOT, "tests", "golden", "Al"), "test_Al")
        mix = {}
        for name, kw in (("pulay", {}), ("response", {"mixing": "response"})):
            sc = SCF(A)
            t0 = time.perf_counter()
            r = sc.run(scf_tol=1e-6, **kw)
            mix[name] = {"iterations": r["iterations"], "converged": r["converged"],
                         "response_products": r["response_products"], "s": time.perf_counter() - t0,
                         "free_energy": r["free_energy"]}
            sc.close()
        log("bench: NVE steps on the Al test input")
        from sqdft_amd.md import NVE
        sc = SCF(A)
        dyn = NVE(sc, masses=[26.9815385], time_step=1.0)
        recs = dyn.run(4)
        sc.close()
        out["md"] = {"workload": "reference tests/Al input, NVE 1 fs, velocities at the electronic T, "
                                 "charge extrapolation, scf_tol 1e-4",
                     "s_per_step": [r["seconds"] for r in recs], "scf_iterations": [r["scf_iterations"] for r in recs],
                     "TE": [r["TE"] for r in recs]}
        A.close()
        out["scf_mixing"] = {"workload": "reference tests/Al input (4 atoms, 10^3 nodes, npl 28), scf_tol 1e-6",
                             **mix}
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(S, veff, cfg["Rcut"], cfg["T"], args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
