#!/usr/bin/env python3
"""Batched GFN1-xTB SCC energy + gradient throughput on one MI355X.

Workload: 4096 independent neutral closed-shell clusters of 2-40 rigid
molecules drawn from {H2O, CH4, NH3} at 0.9 g/cm^3 (instance i seeded with
1000 + i), energy plus full analytic gradient per instance, 300 K Fermi
smearing, SCC tolerances 1e-7 Eh (energy) and 2e-5 (charge rms).

One step = one full energy + gradient pass over the whole batch; inputs are
uploaded to HBM before the timed region.  Two paths are timed on the same batch:
  plain    FP64 throughout, standard (zero) SCC start
  learned  SCC start from the GPU-trained charge model (gfn1-rs_amd/data/guess_mlp.bin)
           + mixed precision (FP32 MFMA density/energy-weighted-density products
           and FP32 gradient contraction; eigensolve, mixing, energy sums FP64)
`value` is the learned + mixed throughput; the plain path and the ratio are
reported beside it.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gfn1-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

# The batch engine runs one HIP stream per matrix-size bucket (14 for this
# workload); the HIP runtime maps streams onto GPU_MAX_HW_QUEUES hardware queues
# and streams that share a queue run one after another.  The run uses 16 queues
# (override with --hw-queues); the variable must be set before the HIP runtime
# initialises, i.e. before the library or torch touches the GPU.
HW_QUEUES_DEFAULT = 16


def _hw_queues_from_argv():
    for i, a in enumerate(sys.argv):
        if a == "--hw-queues" and i + 1 < len(sys.argv):
            return sys.argv[i + 1]
        if a.startswith("--hw-queues="):
            return a.split("=", 1)[1]
    return str(HW_QUEUES_DEFAULT)


os.environ["GPU_MAX_HW_QUEUES"] = _hw_queues_from_argv()

METRIC = "SCC energy+gradient molecules/sec (batched, ML guess+mixed vs plain FP64)"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_tridiag_traffic_batch4096.txt")
MFMA_FILE = os.path.join(ROOT, "profiles", "mfma_fp64_batch4096.txt")


def measured_traffic():
    """HBM bytes per k_tridiag launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_mfma():
    """FP64 matrix-core rate per MFMA kernel from the committed rocprofv3 --pmc
    summary (SQ_INSTS_VALU_MFMA_MOPS_F64 x 512 over the kernel's duration, and
    the SQ_VALU_MFMA_BUSY_CYCLES share), as {kernel: {...}}; {} when absent."""
    out = {}
    try:
        with open(MFMA_FILE) as f:
            rows = f.read().splitlines()
    except OSError:
        return out
    for line in rows:
        parts = line.split()
        if len(parts) == 7 and parts[0].startswith("k_") and parts[5].endswith("%"):
            out[parts[0]] = {"calls": int(parts[1]), "ms": float(parts[2]), "tflops": float(parts[4]),
                             "frac_of_peak": float(parts[5].rstrip("%")) / 100.0,
                             "busy": float(parts[6].rstrip("%")) / 100.0}
    return out


def dense_flops(n, iters):
    """Algorithmic dense FP64 flops of one converged instance (19 n^3 per SCC
    iteration + 11 n^3 setup: congruence 4, eigensolve 9, back-transform 2,
    P/W 4; setup = S factorisation counted as an eigensolve plus 2)."""
    return iters * 19.0 * n ** 3 + 11.0 * n ** 3


def calibrate_copy():
    """Plain streaming kernel of known size for scaling the FETCH_SIZE / WRITE_SIZE
    counters: y = 1.0 * x over 2^27 doubles (1 GiB read, 1 GiB written)."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "elementwise mul", "bytes_read": x.numel() * 8,
                      "bytes_written": y.numel() * 8, "launches": 3}))


def run_path(eng, label, steps, warmup, tune=False):
    for w in range(warmup):
        eng.run()
        print("%s warmup %d: %s" % (label, w, eng.timings()), file=sys.stderr, flush=True)
    tuning = None
    if tune:
        # widths of the back-transformation per size bucket, timed on the
        # warm-up's data (results do not depend on them); one more untimed pass
        # captures the iteration graphs with the chosen widths
        tuning = eng.autotune()
        eng.run()
        print("%s autotuned: %s" % (label, [(t["ld"], t["cols"]) for t in tuning]), file=sys.stderr, flush=True)
    times = []
    tri = [0.0, 0, 0.0]
    for s in range(steps):
        t = time.perf_counter()
        eng.run()
        times.append(time.perf_counter() - t)
        k_ms, k_n, k_b = eng.kernel_stats()
        tri = [tri[0] + k_ms, tri[1] + k_n, tri[2] + k_b]
        print("%s step %d: %.3fs %s" % (label, s, times[-1], eng.timings()), file=sys.stderr, flush=True)
    n = eng.n
    return {"ms": 1000.0 * float(np.mean(times)), "tri": tri, "launch": eng.launch_stats(), "tuning": tuning,
            "status": np.array([eng.status(i) for i in range(n)]),
            "iters": np.array([eng.iterations(i) for i in range(n)]), "phase_ms": eng.timings()}


def standalone_kernel_pass(eng):
    """One more pass of the headline engine with every bucket's work alone on
    the device (GFN1_SERIAL_BUCKETS): the k_tridiag launch durations then
    belong to the kernel by itself, not to a GPU shared with the other
    buckets' kernels.  Not part of the timed steps."""
    os.environ["GFN1_SERIAL_BUCKETS"] = "1"
    try:
        t = time.perf_counter()
        eng.run()
        wall = time.perf_counter() - t
        return eng.kernel_stats(), wall
    finally:
        del os.environ["GFN1_SERIAL_BUCKETS"]


def host_description():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "cores_available": len(os.sched_getaffinity(0))}


def parity(eng, systems, count):
    import oracle
    de, dg = 0.0, 0.0
    for i in range(count):
        z, x = systems[i]
        ref = oracle.run(z, x, energy_tolerance=1e-7)
        de = max(de, abs(ref["energy"] - eng.energy(i)))
        dg = max(dg, float(np.abs(ref["gradient"] - eng.gradient(i)).max()))
    return {"instances": count, "max_abs_dE": de, "max_abs_dg": dg}


SINGLE_WORKLOADS = {
    # name: (molecules, cube side in Angstrom, seed)
    "water48": (48, 11.3, 0),
    "water1000": (1000, 31.0, 1),
    "water2000": (2000, 39.1, 4),
}


def run_single(name, precisions):
    """Standalone single-system workload: prints one JSON line."""
    print(json.dumps(single_system(name, precisions)))


def single_system(name, precisions, oracle_max_ao=1000):
    """One water cluster (energy + gradient, 300 K, SCC tolerance 1e-7 Eh)
    through the same engine: wall time per energy+gradient, iterations and the
    size-independent checks (translation invariance of the gradient, total
    charge); the mixed path is compared with the FP64 path on the same system.
    Prints one JSON line (not the batched headline metric)."""
    from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions, basis_counts, geometry
    nmol, side, seed = SINGLE_WORKLOADS[name]
    t0 = time.time()
    z, x = geometry.water_cluster(nmol, side, seed)
    nsh = basis_counts(z)[0]
    nao = geometry.basis_size(z)[1]
    print("%s: %d atoms, %d AO, generated in %.1fs" % (name, len(z), nao, time.time() - t0), file=sys.stderr,
          flush=True)
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(energy_tolerance=1e-7))
    # kernels and libraries loaded outside the timed run
    BatchEngine([geometry.water_cluster(48, 11.3, seed=0)], options=opts).run()
    os.environ["GFN1_PROGRESS"] = "1"
    out = {"workload": "%s: %d rigid waters in a %.1f A cube (seed %d), %d atoms, %d AO, energy + gradient, "
                       "300 K, SCC energy tolerance 1e-7" % (name, nmol, side, seed, len(z), nao),
           "nao": nao, "natoms": len(z)}
    ref = None
    for prec in precisions:
        eng = BatchEngine([(z, x)], options=opts, precision=prec)
        t = time.perf_counter()
        eng.run()
        dt = time.perf_counter() - t
        g = eng.gradient(0)
        q = eng.shell_charges(0, nsh)
        it = eng.iterations(0)
        r = {"seconds": dt, "status": eng.status(0), "iterations": it, "energy": eng.energy(0),
             "phase_ms": eng.timings(), "sum_gradient_max": float(np.abs(g.sum(0)).max()),
             "charge_sum_error": float(abs(q.sum())),
             "dense_fp64_tflops": dense_flops(nao, it) / dt / 1e12}
        if ref is None:
            ref = (r["energy"], g)
            ref_g = g
        else:
            r["max_abs_dE_vs_fp64"] = abs(r["energy"] - ref[0])
            r["max_abs_dg_vs_fp64"] = float(np.abs(g - ref[1]).max())
        out[prec] = r
        print("%s %s: %.2f s, %d iterations" % (name, prec, dt, it), file=sys.stderr, flush=True)
    if nao <= oracle_max_ao:
        # small enough for the host oracle (one core): its time and the parity of the FP64 result
        import oracle
        t = time.perf_counter()
        ref_o = oracle.run(z, x, energy_tolerance=1e-7)
        out["cpu_baseline"] = {"seconds": time.perf_counter() - t, "cores": 1, "kind": "port",
                               "sample": "the same system, FP64 C++ oracle"}
        if "fp64" in out:
            out["fp64"]["abs_dE_vs_oracle"] = abs(out["fp64"]["energy"] - ref_o["energy"])
            out["fp64"]["max_abs_dg_vs_oracle"] = float(np.abs(ref_g - ref_o["gradient"]).max())
    os.environ.pop("GFN1_PROGRESS", None)
    return out


def periodic_boxes(args):
    """The periodic workload: water boxes of 32 molecules in a 9.86 A cube
    (1.0 g/cm^3), box i seeded with 5000 + i, every box its own periodic cell."""
    from gfn1_amd import geometry
    return [geometry.water_box(32, 9.86, 5000 + i) for i in range(args.periodic_boxes)]


def periodic_workload(args):
    """Batched periodic Gamma-point energy + gradient (pbc/gradient.rs:64-73 per
    box) through the same engine: boxes per second over the batch, FP64, SCC
    energy tolerance 1e-7; parity of the first boxes against the oracle's
    periodic restatement, whose single-core time on one box is the CPU
    baseline.  Returns a dict (an extra key of the default run)."""
    from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions
    t0 = time.time()
    boxes = periodic_boxes(args)
    print("periodic: %d boxes generated in %.1fs" % (len(boxes), time.time() - t0), file=sys.stderr, flush=True)
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(energy_tolerance=1e-7))
    eng = BatchEngine([(z, x, 0.0, c) for z, x, c in boxes], options=opts)
    eng.run()  # warm-up: kernels loaded, graphs captured
    times = []
    for s in range(max(args.steps, 1)):
        t = time.perf_counter()
        eng.run()
        times.append(time.perf_counter() - t)
        print("periodic step %d: %.3fs %s" % (s, times[-1], eng.timings()), file=sys.stderr, flush=True)
    ms = 1000.0 * float(np.mean(times))
    it = np.array([eng.iterations(i) for i in range(len(boxes))])
    st = np.array([eng.status(i) for i in range(len(boxes))])
    out = {"workload": "%d periodic boxes of 32 rigid waters in a 9.86 A cube (96 atoms, 256 AO, seeds 5000+i), "
                       "Gamma point, energy + gradient, FP64, 300 K, SCC energy tolerance 1e-7" % len(boxes),
           "boxes_per_s": len(boxes) / (ms / 1000.0), "ms_per_step": ms, "steps": len(times),
           "mean_scc_iterations": float(it.mean()), "converged": int((st == 0).sum()), "phase_ms": eng.timings()}
    if args.check > 0:
        import oracle
        de, dg = 0.0, 0.0
        t = time.perf_counter()
        for i in range(min(args.check, 2)):
            z, x, c = boxes[i]
            ref = oracle.run(z, x, cell=c, energy_tolerance=1e-7)
            if i == 0:
                cpu_s = time.perf_counter() - t
            de = max(de, abs(ref["energy"] - eng.energy(i)))
            dg = max(dg, float(np.abs(ref["gradient"] - eng.gradient(i)).max()))
        out["parity_vs_oracle"] = {"boxes": min(args.check, 2), "max_abs_dE": de, "max_abs_dg": dg}
        out["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "boxes/s", "cores": 1, "kind": "port",
                               "sample": "box 0, FP64 C++ oracle (periodic restatement), one core"}
    return out


def run_optimize(args):
    """Batched L-BFGS geometry optimisation (optimize_geometry of every
    cluster of the batch at once, default options except max_iterations)
    through the same engine: accepted L-BFGS steps per second over the whole
    batch, on the plain FP64 path and on the learned-start + mixed-precision
    path with warm-started SCCs.  The CPU baseline is the oracle's L-BFGS
    restatement on the first clusters, one core.  Prints one JSON line (not
    the batched headline metric)."""
    from gfn1_amd import (AnalyticGradientOptions, BatchEngine, ElectronicOptions, GeometryOptimizationOptions,
                          GuessModel, geometry)
    systems = geometry.mixed_batch(args.batch, first_seed=1000)
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(energy_tolerance=1e-7))
    gopt = GeometryOptimizationOptions(max_iterations=args.opt_iterations, keep_trajectory=False)
    model = GuessModel.load(args.model)
    out = {"metric": "L-BFGS steps/s", "unit": "accepted steps/s", "higher_is_better": True, "n_gpus": 1,
           "data": "synthetic",
           "config": {"workload": "batched L-BFGS relaxation of mixed H2O/CH4/NH3 clusters, 2-40 molecules, "
                                  "0.9 g/cm3, seeds 1000+i, SCC energy tolerance 1e-7, max_iterations %d"
                                  % args.opt_iterations, "global_batch": len(systems)},
           "paths": {}}
    for label, kw, warm in (("plain", {}, False),
                            ("learned+mixed+warm", {"guess": model, "precision": "mixed"}, True)):
        eng = BatchEngine(systems, options=opts, **kw)
        eng.optimize(GeometryOptimizationOptions(max_iterations=1, keep_trajectory=False))  # untimed warm-up
        eng.set_positions([x for _, x in systems])
        gopt.warm_start = warm
        t = time.perf_counter()
        res = eng.optimize(gopt)
        dt = time.perf_counter() - t
        st = eng.optimization_stats
        ok = [r for r in res if r.status == 0]
        steps = sum(r.iterations for r in ok)
        out["paths"][label] = {
            "steps_per_s": steps / dt, "seconds": dt, "accepted_steps": steps, "evaluations": st["evaluations"],
            "evaluations_per_s": st["evaluations"] / dt, "rounds": st["rounds"],
            "converged": sum(1 for r in ok if r.converged), "failed": len(res) - len(ok),
            "mean_final_max_gradient": float(np.mean([r.max_gradient for r in ok])) if ok else None}
        print("optimize %s: %.2f s, %d steps, %d evaluations in %d rounds, %d converged" % (
            label, dt, steps, st["evaluations"], st["rounds"], out["paths"][label]["converged"]),
            file=sys.stderr, flush=True)
    head = out["paths"]["learned+mixed+warm"]
    out["value"] = head["steps_per_s"]
    if args.opt_cpu_sample > 0:
        import optimizer_oracle
        n = min(args.opt_cpu_sample, len(systems))
        t = time.perf_counter()
        steps = 0
        for z, x in systems[:n]:
            r = optimizer_oracle.optimize_gfn1(z, x, scc_options={"energy_tolerance": 1e-7},
                                               max_iterations=args.opt_iterations)
            steps += r["iterations"]
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": steps / dt, "unit": "accepted steps/s", "cores": 1, "kind": "port",
                               "sample": "first %d clusters, oracle L-BFGS over the FP64 C++ oracle" % n}
    print(json.dumps(out))


def run_md_workload(args):
    """Batched NVE molecular dynamics (velocity Verlet, 0.5 fs, 300 K
    Maxwell-Boltzmann start velocities) of the clusters of the batch, FP64,
    each SCC warm-started from the previous step's charges: system-steps per
    second.  Prints one JSON line (not the batched headline metric)."""
    from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions, geometry
    from gfn1_amd.md import run_md
    systems = geometry.mixed_batch(args.batch, first_seed=1000)
    eng = BatchEngine(systems, options=AnalyticGradientOptions(electronic=ElectronicOptions(energy_tolerance=1e-7)))
    eng.run()  # untimed warm-up
    t = time.perf_counter()
    res = run_md(eng, [x for _, x in systems], args.md_steps, timestep_fs=0.5, temperature=300.0, seed=1)
    dt = time.perf_counter() - t
    ok = [i for i, s in enumerate(res.status) if s == 0]
    drifts = np.abs(res.total[:, ok] - res.total[0, ok]).max(axis=0) if ok else None
    out = {"metric": "MD system-steps/s", "unit": "system-steps/s", "higher_is_better": True, "n_gpus": 1,
           "data": "synthetic", "value": len(ok) * args.md_steps / dt, "seconds": dt, "steps": args.md_steps,
           "systems": len(systems), "failed": len(systems) - len(ok),
           "total_energy_drift": None if drifts is None else {"median": float(np.median(drifts)),
                                                               "max": float(drifts.max())},
           "config": {"workload": "NVE velocity Verlet of mixed H2O/CH4/NH3 clusters (seeds 1000+i, 2-40 molecules), "
                                  "0.5 fs, 300 K start, FP64, SCC energy tolerance 1e-7, warm-started SCCs",
                      "global_batch": len(systems)}}
    print("md: %d systems x %d steps in %.2f s" % (len(systems), args.md_steps, dt), file=sys.stderr, flush=True)
    print(json.dumps(out))


def run_hessian(args):
    """Cartesian Hessians of the first clusters of the batch (central
    differences of the analytic gradient, step 1e-4 bohr, SCC tolerances
    1e-11 / 1e-11, displaced SCCs started from the undisplaced charges):
    Hessians/s and displaced-geometry evaluations/s.  The CPU
    baseline is the oracle's finite-difference Hessian of the smallest of
    those clusters, one core.  Prints one JSON line (not the batched headline
    metric)."""
    from gfn1_amd import geometry, hessian, hessian_batch
    systems = geometry.mixed_batch(args.hess_systems, first_seed=1000)
    nat = [len(z) for z, _ in systems]
    small = min(range(len(systems)), key=lambda i: nat[i])
    hessian(*systems[small])  # untimed warm-up
    t = time.perf_counter()
    hs, status = hessian_batch(systems)
    dt = time.perf_counter() - t
    evals = sum(1 + 6 * n for n in nat)
    out = {"metric": "Hessians/s", "unit": "Hessians/s", "higher_is_better": True, "n_gpus": 1, "data": "synthetic",
           "value": len(systems) / dt, "seconds": dt, "failed": sum(1 for s in status if s != 0),
           "gradient_evaluations": evals, "evaluations_per_s": evals / dt,
           "config": {"workload": "Cartesian Hessians of mixed H2O/CH4/NH3 clusters (seeds 1000+i, 2-40 molecules), "
                                  "central differences of the analytic gradient, step 1e-4 bohr, SCC tolerances "
                                  "1e-11 / 1e-11, warm-started displaced SCCs, FP64", "systems": len(systems), "atoms_min": min(nat),
                      "atoms_max": max(nat), "atoms_mean": float(np.mean(nat))}}
    print("hessian: %d systems, %d evaluations in %.2f s" % (len(systems), evals, dt), file=sys.stderr, flush=True)
    if args.opt_cpu_sample > 0:
        import hessian_oracle
        z, x = systems[small]
        t = time.perf_counter()
        hessian_oracle.fd_hessian(z, x)
        dtc = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 1.0 / dtc, "unit": "Hessians/s", "cores": 1, "kind": "port",
                               "sample": "smallest cluster (%d atoms), oracle finite-difference Hessian" % nat[small],
                               "evaluations_per_s": (1 + 6 * nat[small]) / dtc}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--eigensolver", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=256, help="instances timed on the host oracle (0 = skip)")
    ap.add_argument("--check", type=int, default=4, help="instances checked against the oracle")
    ap.add_argument("--path", choices=("both", "plain", "learned"), default="both",
                    help="paths to time (profiling runs use one)")
    ap.add_argument("--model", default=os.path.join(ROOT, "gfn1-rs_amd", "data", "guess_mlp.bin"))
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB read + 1 GiB write elementwise kernel (HBM counter calibration)")
    ap.add_argument("--opt-iterations", type=int, default=20, help="max_iterations of the optimize workload")
    ap.add_argument("--opt-cpu-sample", type=int, default=1,
                    help="clusters optimised by the host oracle for the optimize workload (0 = skip)")
    ap.add_argument("--hess-systems", type=int, default=32, help="clusters of the hessian workload")
    ap.add_argument("--md-steps", type=int, default=10, help="steps of the md workload")
    ap.add_argument("--periodic-boxes", type=int, default=512, help="water boxes of the periodic workload")
    ap.add_argument("--workload", choices=["batch", "optimize", "hessian", "md", "periodic"] + sorted(SINGLE_WORKLOADS),
                    default="batch",
                    help="batch: the batched headline metric; waterN: one large cluster (separate report)")
    ap.add_argument("--precision", choices=("both", "fp64", "mixed"), default="both",
                    help="precisions run for a single-system workload (fp64 first)")
    ap.add_argument("--extras", type=int, default=1,
                    help="1: also time the single systems (384, 8000 and 16000 AO) and the learned-start protocol "
                         "(20000 training / 4096 evaluation snapshots) after the headline, as extra keys")
    ap.add_argument("--standalone-pass", type=int, default=1,
                    help="1: one extra pass with every size bucket alone on the device, for the roofline figures "
                         "(0 only for profiles of the concurrent pass)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: autotune the back-transformation width per size bucket after the warm-up passes")
    ap.add_argument("--compare-host-loop", type=int, default=1,
                    help="1: time one extra pass of the headline path with host-launched SCC iterations "
                         "(no graph replays), reported under scc_launch")
    ap.add_argument("--hw-queues", type=int, default=HW_QUEUES_DEFAULT, choices=range(1, 33), metavar="1..32",
                    help="GPU_MAX_HW_QUEUES for this run (set before the HIP runtime starts)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "periodic":
        if args.gpus != 1:
            print("bench.py measures one GPU (--gpus %d refused)" % args.gpus, file=sys.stderr)
            sys.exit(2)
        print(json.dumps(periodic_workload(args)))
        return
    if args.workload in ("optimize", "hessian", "md"):
        if args.gpus != 1:
            print("bench.py measures one GPU (--gpus %d refused)" % args.gpus, file=sys.stderr)
            sys.exit(2)
        {"optimize": run_optimize, "hessian": run_hessian, "md": run_md_workload}[args.workload](args)
        return
    if args.workload != "batch":
        if args.gpus != 1:
            print("bench.py measures one GPU (--gpus %d refused)" % args.gpus, file=sys.stderr)
            sys.exit(2)
        run_single(args.workload, ["fp64", "mixed"] if args.precision == "both" else [args.precision])
        return
    if args.gpus != 1:
        print("bench.py measures one GPU; every run here has exactly one device (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from gfn1_amd import AnalyticGradientOptions, BatchEngine, ElectronicOptions, GuessModel, geometry

    t0 = time.time()
    systems = geometry.mixed_batch(args.batch, first_seed=1000)
    nao = np.array([geometry.basis_size(z)[1] for z, _ in systems])
    print("generated %d clusters in %.1fs (AO %d..%d, mean %.1f)" % (len(systems), time.time() - t0, nao.min(),
                                                                   nao.max(), nao.mean()), file=sys.stderr, flush=True)
    opts = AnalyticGradientOptions(electronic=ElectronicOptions(energy_tolerance=1e-7))
    res, engines = {}, {}
    if args.path in ("both", "plain"):
        engines["plain"] = BatchEngine(systems, options=opts, eigensolver=args.eigensolver)
        res["plain"] = run_path(engines["plain"], "plain", args.steps, args.warmup, bool(args.autotune))
    if args.path in ("both", "learned"):
        model = GuessModel.load(args.model)
        engines["learned"] = BatchEngine(systems, options=opts, eigensolver=args.eigensolver, guess=model,
                                         precision="mixed")
        res["learned"] = run_path(engines["learned"], "learned+mixed", args.steps, args.warmup, bool(args.autotune))
    head = "learned" if "learned" in res else "plain"
    r = res[head]
    ms = r["ms"]
    value = len(systems) / (ms / 1000.0)
    flops = float(sum(dense_flops(n, it) for n, it in zip(nao, r["iters"])))
    dense_tflops = flops / (ms / 1000.0) / 1e12
    # concurrent launches (timed steps): durations include sharing the GPU
    # (buckets iterating by graph replays record no per-launch events)
    tri_ms_c, tri_n_c, tri_bytes_c = r["tri"]
    conc_avg_ms = tri_ms_c / tri_n_c if tri_n_c else None
    # the same kernel alone on the device (one extra, untimed pass)
    scc_launch = {"per_pass": r["launch"],
                  "note": "SCC phase of one headline pass: kernels run, host enqueue calls (one per kernel launch, "
                          "library call or graph replay), graph replays, host-launched iterations, host reads of a "
                          "bucket's gathered states"}
    if args.compare_host_loop:
        eng_h = BatchEngine(systems, options=opts, eigensolver=args.eigensolver, scc_graphs=False,
                            **({"guess": GuessModel.load(args.model), "precision": "mixed"} if head == "learned" else {}))
        rh = run_path(eng_h, "host loop", 1, 1, bool(args.autotune))
        scc_launch["host_loop"] = {"ms_per_step": rh["ms"], "per_pass": rh["launch"],
                                   "same_iterations": bool(np.array_equal(rh["iters"], r["iters"]))}
        del eng_h
    if args.standalone_pass:
        (tri_ms, tri_n, tri_bytes), serial_wall = standalone_kernel_pass(engines[head])
    else:  # profiling runs of the concurrent pass only: no roofline figures
        (tri_ms, tri_n, tri_bytes), serial_wall = (0.0, 0, 0.0), 0.0
    tri_avg_ms = tri_ms / max(tri_n, 1)
    tri_bytes_per_launch = tri_bytes / max(tri_n, 1)
    achieved = tri_bytes_per_launch / (tri_avg_ms / 1000.0) / 1e9 if tri_n else 0.0

    paths = {}
    for name, rr in res.items():
        paths[name] = {"molecules_per_s": len(systems) / (rr["ms"] / 1000.0), "ms_per_step": rr["ms"],
                       "mean_scc_iterations": float(rr["iters"].mean()),
                       "converged": int((rr["status"] == 0).sum()), "phase_ms": rr["phase_ms"],
                       "parity_vs_oracle": parity(engines[name], systems, args.check) if args.check > 0 else None}
    if "plain" in res and "learned" in res:
        ep, el = engines["plain"], engines["learned"]
        paths["learned"]["max_abs_dE_vs_plain"] = float(max(abs(ep.energy(i) - el.energy(i))
                                                            for i in range(len(systems))))
        paths["learned"]["speedup_vs_plain"] = res["plain"]["ms"] / res["learned"]["ms"]

    cpu = None
    if args.cpu_sample > 0:
        import oracle
        # every host core this process may run on (16 on a one-GPU machine)
        nthreads = len(os.sched_getaffinity(0)) or 1
        sample = systems[: args.cpu_sample]
        t = time.perf_counter()
        oracle.run_batch(sample, nthreads=nthreads, energy_tolerance=1e-7)
        dt = time.perf_counter() - t
        cpu = {"value": len(sample) / dt, "unit": "molecules/s", "cores": nthreads, "kind": "port",
               "sample": "first %d clusters of the batch (%.1f s), FP64 C++ oracle (Lowdin orthogonaliser, "
                         "Householder + QL eigensolver, full P/W products), OpenMP over instances, one thread "
                         "per instance" % (len(sample), dt)}
        cpu.update(host_description())

    out = {
        "metric": METRIC, "value": value, "unit": "molecules/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64/f32" if head == "learned" else "f64", "data": "synthetic",
        "config": {"workload": "batched mixed H2O/CH4/NH3 clusters, 2-40 molecules, 0.9 g/cm3, seeds 1000+i, "
                               "energy + gradient, 300 K, SCC energy tolerance 1e-7",
                   "global_batch": len(systems), "ao_min": int(nao.min()), "ao_max": int(nao.max()),
                   "ao_mean": float(nao.mean()),
                   "path": "learned start + mixed precision" if head == "learned" else "plain FP64",
                   "parallelism": "single GPU", "hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"])},
        "paths": paths,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_tridiag", "launches": tri_n, "avg_launch_ms": tri_avg_ms,
                     "algorithmic_bytes_per_launch": tri_bytes_per_launch,
                     "kernel_ms_per_pass_alone": tri_ms, "serial_pass_ms": 1000.0 * serial_wall,
                     "concurrent_avg_launch_ms": conc_avg_ms,
                     "limiter": "latency: one work-group per matrix runs n dependent column steps (a symmetric "
                                "matrix-vector product, a reflector and a rank-2 update each, joined by "
                                "work-group barriers); launches over fewer than 64 matrices move about "
                                "0.02 TB/s, full launches 2.5-2.9 TB/s (profiles/k_tridiag_traffic_batch4096.txt)",
                     "note": "algorithmic bytes = 8 (n-1) n (n+1) / 6 per matrix (one pass over the trailing "
                             "lower triangle per column step), summed over the launch; launch durations from HIP "
                             "events on the bucket stream in one extra pass of the headline path with every "
                             "bucket alone on the device (during the timed steps the buckets overlap and a "
                             "launch shares the GPU: concurrent_avg_launch_ms)"},
        "dense_fp64": {"achieved_tflops": dense_tflops, "peak_tflops": FP64_PEAK_TFLOPS,
                       "frac": dense_tflops / FP64_PEAK_TFLOPS,
                       "note": "whole-pass dense work 19 n^3 per SCC iteration + 11 n^3 over wall time, headline path"},
        "scc_launch": scc_launch,
        "autotune": {"kernel": "k_backtransform", "choices": r["tuning"],
                     "note": "eigenvectors per work-group per size bucket (ld = padded dimension), mean launch ms "
                             "per width timed on the warm-up pass's reflectors; every width gives identical bits"},
        "mfma_fp64": {"peak_tflops": FP64_PEAK_TFLOPS, "source": "profiles/mfma_fp64_batch4096.txt",
                      "kernels": measured_mfma()},
        "cpu_baseline": cpu, "phase_ms": engines[head].timings(),
    }
    if args.extras:
        del engines
        out["single_systems"] = {}
        for name in ("water48", "water1000", "water2000"):
            print("single system %s" % name, file=sys.stderr, flush=True)
            out["single_systems"][name] = single_system(name, ["fp64", "mixed"], oracle_max_ao=400)
        print("learned-start protocol", file=sys.stderr, flush=True)
        from gfn1_amd import train_guess
        out["learned_start_protocol"] = train_guess.protocol()
        print("periodic workload", file=sys.stderr, flush=True)
        out["periodic"] = periodic_workload(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
