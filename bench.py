"""Benchmark of the SQ hot path (Clenshaw-Curtis spectral quadrature).

A step is one pass of the per-node SQ loop (sq.cpp:530-661: Lanczos spectral
bounds and Chebyshev moments of every node) over one batch of FD nodes, the
unit the large workload is launched in: 8,192 nodes of the 864-atom FCC Al
cell (96^3 FD nodes, 12th-order FD, nloc = 17, 35^3 cube, npl = 48,
T = 58,000 K).  Batches are taken in the library's visiting order and cycle
through the cell.  Veff and the projector tables are resident on the device;
each step uploads Veff (7 MB) and returns the batch's moments, and its
kernels dominate (99 %).  The same batch steps in mixed precision (FP32
storage) and with FP32 Lanczos passes only.  Then, once each: the whole
density step of the cell (every batch, the Fermi level, rho) in FP64, mixed
precision and FP32 Lanczos, with the density errors against FP64; an SCF
iteration and the forces of the 216-atom
LiH cell; SCF mixing and NVE steps on the reference's Al test input; the
FP64 oracle on the host's cores.  Prints one JSON line.
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
    "al864": ("FCC Al 6x6x6 (864 atoms), 96^3 nodes, FD order 12, nloc=17 (35^3 cube), npl=48, "
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


KERNELS = [("k_zmarch<CHEB>", "Chebyshev recurrence: read t_k, t_{k-1}, write t_{k+1}", 3),
           ("k_zmarch<LANCZOS>", "Lanczos pass: read r', v_{k-1}, write v_k, r'_{k+1}", 4),
           ("projections<LANCZOS>", "projections of the Lanczos vector: read r', v_{k-1}", 2),
           ("projections<CHEB>", "projections of t_k: read t_k", 1)]


def roofline(sq, workload, precision, N):
    """The hot kernels of the last call, timed with HIP events around every
    launch on the library's stream; the dominant one is the roofline."""
    rs = 8 if precision == "fp64" else 4
    traffic = profiled_traffic(workload, precision)
    rl = []
    for w, (key, what, nvec) in enumerate(KERNELS):
        k = sq.kernel_timing(w)
        avg_ms = k["total_ms"] / max(k["launches"], 1)
        alg = k["alg_bytes"] / max(k["launches"], 1)
        rl.append({"kernel": f"{key} ({what})", "launches": k["launches"], "avg_ms": avg_ms,
                   "total_ms": k["total_ms"], "alg_bytes_per_launch": alg,
                   "achieved_GBs": alg / (avg_ms * 1e-3) / 1e9 if avg_ms else 0.0,
                   "alg_bytes_per_unit": f"{nvec} * {N} * {rs} B per node-step",
                   "traffic": traffic.get(key)})
    dom = int(np.argmax([r["total_ms"] for r in rl]))
    d = rl[dom]
    return {"bound": "hbm", "kernel": d["kernel"], "achieved": d["achieved_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": d["achieved_GBs"] / HBM_PEAK_GBS, "traffic": d["traffic"],
            "alg_bytes_per_unit": d["alg_bytes_per_unit"], "launches": d["launches"], "avg_ms": d["avg_ms"],
            "traffic_source": (f"rocprofv3 FETCH_SIZE / WRITE_SIZE per dispatch, scaled by a plain copy "
                               f"(profiles/traffic_{workload}.md)") if d["traffic"] else None,
            "others": [r for i, r in enumerate(rl) if i != dom]}


def batch_steps(sq, veff, order, B, warmup, steps, label):
    """Times `steps` batch steps after `warmup` untimed ones; returns
    (seconds per step, nodes per step, last result)."""
    nbat = (order.size + B - 1) // B
    nodes_of = lambda i: order[(i % nbat) * B:(i % nbat) * B + B]
    for i in range(warmup):
        sq.node_moments(veff, nodes_of(i))
        log(f"bench: {label} warmup {i}")
    times, count, last = [], 0, None
    for i in range(warmup, warmup + steps):
        nd = nodes_of(i)
        t0 = time.perf_counter()
        last = sq.node_moments(veff, nd)
        times.append(time.perf_counter() - t0)
        count += nd.size
        log(f"bench: {label} step {i - warmup} {times[-1]:.3f} s (Lanczos {last['iters'].min()}..{last['iters'].max()})")
    return float(np.sum(times)) / steps, count / steps, last


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="al864", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--batch-nodes", type=int, default=8192)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full", action="store_true", help="skip the whole-cell density steps")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--mixing-workload", default="test_Al", choices=["test_Al"] + sorted(WORKLOADS),
                    help="cell of the SCF mixing comparison (Periodic Pulay vs response mixing to 1e-6)")
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
    from sqdft_amd.sq import SpectralQuadrature, node_order

    cfg = configs.CONFIGS[args.workload]
    log(f"bench: building {args.workload}")
    S = configs.build(args.workload, seed=0)
    veff = configs.smooth_veff(S)
    info = S.info
    nn = S.nn
    B = min(args.batch_nodes if args.batch_nodes > 0 else nn, nn)
    order = node_order(info["n"])
    N = (2 * info["nloc"] + 1) ** 3
    sq = SpectralQuadrature(S, precision=args.precision, batch_nodes=B)
    if args.copy_calib_mib > 0:
        calibration_copy(args.copy_calib_mib * (1 << 20))
    sec, per, last = batch_steps(sq, veff, order, B, args.warmup, args.steps, args.precision)
    value = per / sec
    out = {"metric": METRIC, "value": value, "unit": "nodes/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * sec, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
           "data": "synthetic (seeded atomic positions; smooth model Veff)",
           "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}", "nodes": nn,
                      "nloc": info["nloc"], "npl": info["npl"], "precision": args.precision,
                      "batch_nodes": B, "nodes_per_step": per,
                      "step": "Lanczos bounds + Chebyshev moments of one batch (sqg_node_moments)"},
           "roofline": roofline(sq, args.workload, args.precision, N)}
    sq.close()
    if not args.no_mixed and args.precision == "fp64":
        log("bench: mixed precision batch steps")
        sqm = SpectralQuadrature(S, precision="mixed", batch_nodes=B)
        msec, mper, _ = batch_steps(sqm, veff, order, B, 1, 4, "mixed")
        out["mixed"] = {"value": mper / msec, "unit": "nodes/s", "ms_per_step": 1e3 * msec,
                        "speedup_vs_fp64": sec / msec,
                        "what": "FP32 storage of every vector, Veff and table; FP64 arithmetic and reductions",
                        "roofline": roofline(sqm, args.workload, "mixed", N)}
        sqm.close()
        log("bench: mixed_lanczos batch steps")
        sql = SpectralQuadrature(S, precision="mixed_lanczos", batch_nodes=B)
        lsec, lper, _ = batch_steps(sql, veff, order, B, 1, 4, "mixed_lanczos")
        out["mixed_lanczos"] = {"value": lper / lsec, "unit": "nodes/s", "ms_per_step": 1e3 * lsec,
                                "speedup_vs_fp64": sec / lsec,
                                "what": "the Lanczos passes as in mixed, the Chebyshev moments in FP64"}
        sql.close()
    if not args.no_full:
        # the whole density step (every batch, global bounds, Fermi level, rho,
        # Ebs, Eent); both Fermi levels solved to 1e-12 so that the density
        # difference is the mixed arithmetic's own (at the reference's 1e-6 the
        # two roots alone may differ by ~1e-6)
        log("bench: full fp64 density step")
        r64 = SpectralQuadrature(S, precision="fp64", batch_nodes=B, fermi_tol=1e-12)
        t0 = time.perf_counter()
        d64 = r64.density(veff, scf_iter=1)
        t64 = time.perf_counter() - t0
        st = d64["stats"]
        r64.close()
        out["full_step"] = {"s": t64, "nodes_per_s": nn / t64, "precision": "fp64", "fermi_tol": 1e-12,
                            "lambda_f": d64["lambda_f"], "electrons": float(d64["rho"].sum() * info["h"] ** 3),
                            "phases_s": {"lanczos": st["t_lanczos"], "cheb": st["t_cheb"],
                                         "fermi": st["t_fermi"], "rho": st["t_rho"]},
                            "lanczos_iters": {"node0": st["lanczos_iters_node0"], "min": st["lanczos_iters_min"],
                                              "max": st["lanczos_iters_max"], "passes": st["lanczos_passes"]},
                            "brent_iters": st["fermi_iters"]}
        if not args.no_mixed:
            log("bench: full mixed density step")
            r32 = SpectralQuadrature(S, precision="mixed", batch_nodes=B, fermi_tol=1e-12)
            t0 = time.perf_counter()
            d32 = r32.density(veff, scf_iter=1)
            t32 = time.perf_counter() - t0
            r32.close()
            out["full_step"]["mixed"] = {
                "s": t32, "nodes_per_s": nn / t32, "speedup_vs_fp64": t64 / t32,
                "rho_rel_err_vs_fp64": float(np.max(np.abs(d32["rho"] - d64["rho"])) / np.max(np.abs(d64["rho"]))),
                "rho_rel_err_l2_vs_fp64": float(np.linalg.norm(d32["rho"] - d64["rho"]) / np.linalg.norm(d64["rho"])),
                "lambda_f_err": abs(d32["lambda_f"] - d64["lambda_f"]),
                "ebs_rel_err": abs(d32["ebs"] - d64["ebs"]) / abs(d64["ebs"])}
            log("bench: full mixed_lanczos density step")
            rl = SpectralQuadrature(S, precision="mixed_lanczos", batch_nodes=B, fermi_tol=1e-12)
            t0 = time.perf_counter()
            dl = rl.density(veff, scf_iter=1)
            tl = time.perf_counter() - t0
            rl.close()
            out["full_step"]["mixed_lanczos"] = {
                "s": tl, "nodes_per_s": nn / tl, "speedup_vs_fp64": t64 / tl,
                "rho_rel_err_vs_fp64": float(np.max(np.abs(dl["rho"] - d64["rho"])) / np.max(np.abs(d64["rho"]))),
                "rho_rel_err_l2_vs_fp64": float(np.linalg.norm(dl["rho"] - d64["rho"]) / np.linalg.norm(d64["rho"])),
                "lambda_f_err": abs(dl["lambda_f"] - d64["lambda_f"]),
                "ebs_rel_err": abs(dl["ebs"] - d64["ebs"]) / abs(d64["ebs"])}
    if not args.no_scf:
        # one SCF iteration of the 216-atom LiH cell on its own pseudocharge:
        # SQ density, total energy, AAJ Poisson from the previous phi, LDA
        # potential and the Veff residual (src/scf.cpp:145-200); the second
        # iteration is timed, then the forces at that state, FP64 and mixed
        log("bench: lih216 scf iterations")
        from sqdft_amd.scf import SCF
        L = configs.build("lih216", seed=0)
        scf = SCF(L, precision="fp64")
        t0 = time.perf_counter()
        scf.Pseudocharge()
        t_pc = time.perf_counter() - t0
        scf.initial_potential()
        it1 = scf.iteration()
        scf.veff = it1["veff_new"]
        it2 = scf.iteration()
        log("bench: lih216 forces")
        t0 = time.perf_counter()
        scf.ForcesLocal()
        t_loc = time.perf_counter() - t0
        t0 = time.perf_counter()
        fnl = scf.sq.ForcesNonLocal()
        t_nl = time.perf_counter() - t0
        v2 = scf.veff
        scf.close()
        log("bench: lih216 mixed nonlocal forces at the same Veff")
        sqm = SpectralQuadrature(L, precision="mixed")
        sqm.density(v2, scf_iter=2, lambda_f=it2["lambda_f"])
        fnl32 = sqm.ForcesNonLocal()
        sqm.close()
        out["scf_step"] = {"workload": "lih216: " + WORKLOADS["lih216"],
                           "s": it2["t_sq"] + it2["t_rest"], "sq_s": it2["t_sq"],
                           "energy_poisson_vxc_s": it2["t_rest"], "poisson_iters": it2["poisson"][0],
                           "scf_err": it2["scf_err"], "lambda_f": it2["lambda_f"],
                           "forces_local_s": t_loc, "forces_nonlocal_s": t_nl, "pseudocharge_s": t_pc,
                           "mixed_nonlocal_force_max_err": float(np.max(np.abs(fnl32 - fnl))),
                           "nonlocal_force_max": float(np.max(np.abs(fnl))),
                           "what": "second SCF iteration from the atomic starting density, unmixed Veff; "
                                   "then local and nonlocal forces at that state"}
        # SCF iterations to SCF_err 1e-6 on the reference's Al test input:
        # its Periodic Pulay mixing vs Newton steps from the adjoint response
        from sqdft_amd.system import System
        A = System.from_files(os.path.join(ROOT, "tests", "golden", "Al"), "test_Al")
        if args.mixing_workload == "test_Al":
            log("bench: scf mixing on the Al test input")
            MX, mx_what = A, "reference tests/Al input (4 atoms, 10^3 nodes, npl 28), scf_tol 1e-6"
        else:
            log(f"bench: scf mixing on {args.mixing_workload}")
            MX = configs.build(args.mixing_workload, seed=0)
            mx_what = f"{args.mixing_workload}: {WORKLOADS[args.mixing_workload]}, scf_tol 1e-6, 40 iterations at most"
        mix = {}
        for name, kw in (("pulay", {}), ("response", {"mixing": "response"})):
            sc = SCF(MX)
            t0 = time.perf_counter()
            r = sc.run(scf_tol=1e-6, progress=lambda i, e: log(f"bench: {name} iteration {i} SCF_err {e:.3e}"),
                       **kw)
            mix[name] = {"iterations": r["iterations"], "converged": r["converged"],
                         "response_products": r["response_products"], "s": time.perf_counter() - t0,
                         "free_energy": r["free_energy"],
                         "scf_err": [h["scf_err"] for h in r["history"]]}
            sc.close()
        out["scf_mixing"] = {"workload": mx_what, **mix}
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
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(S, veff, cfg["Rcut"], cfg["T"], args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
