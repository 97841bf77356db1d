"""Benchmark: seconds per SCF iteration (H/S build + eigensolve) on one GPU.

One step is one Mode-B SCF iteration of the OLCAO path (intgPSCF + bandPSCF
semantics): the overlap sweep S(k), the fused Hamiltonian sweep H(k) with
the potential coefficients folded in (kinetic + potential terms + nuclear,
three-centre lattice sums), contraction, Bloch sums, and the batched FP64
generalised eigensolve for every k-point.  Inputs are resident on the
device before the timed region (the system and screened work lists are
built once, as the reference's setup does).

Both forms run: the plain FP64 drop-in and the chosen mixed form (the
three-Gaussian exponentials by an FP64 range reduction and the FP32 hardware
exponential); `value` is the mixed form, the FP64 time and the mixed form's
eigenvalue and band-energy errors against it are reported beside it.

Workload: si64 -- Si64 2x2x2 diamond supercell (a = 5.431 A, 0.05 A seeded
jitter), s+p+d per Si from 24 even-tempered exponents 0.05..1e7, 16
potential Gaussians 0.1..1e5 per site, 2x2x2 Monkhorst-Pack mesh (8 k),
576 orbitals; synthetic seeded data.

The JSON line carries the roofline of the dominant kernel (the three-centre
sweep: algorithmic FP64 operations per launch from olcao_amd.workmodel over
the device's own evaluation counters, divided by its HIP-event time) and a
CPU baseline: the oracle (a C restatement of the reference algorithm) on one
host core over a bounded sample of atom pairs, extrapolated by the
three-centre work ratio, plus LAPACK zhegv for all k.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s per SCF iter (H/S build+eigensolve), mixed vs FP64 GPU at ≤1 meV band error"
WORKLOADS = {
    "si64": "si64: Si64 2x2x2 diamond supercell, s+p+d/Si, 24 exponents 0.05..1e7, "
            "16 potential Gaussians/site, 8 k, n=576",
    "diamond_c8": "diamond_c8: diamond C8, 1s 2s 2p from 20 exponents 0.12..1e6, "
                  "16 potential Gaussians, 4x4x4 k (64 k), n=40",
    "quartz": "quartz: alpha-quartz SiO2 (9 atoms), Si s+p+d, O 1s 2s 2p, 8x8x8 k (512 k), n=57",
    "asio2_1000": "asio2_1000: amorphous SiO2 cube, 333 Si + 667 O at 2.2 g/cm^3, minimal basis "
                  "(Si 1s-3p, O 2s 2p + 1s core), Gamma only (real), n=5665, 667 core states",
    "uneven_batch": "uneven_batch: 64 seeded cells of 8-216 C/Si/O atoms on jittered diamond/"
                    "zincblende lattices at 2.2-3.5 g/cm^3, minimal or extended basis, 1-64 k "
                    "each (n = 40..1296), one batched call per iteration",
}
FP64_PEAK_TFLOPS = 78.6
# HBM bytes of one mixed-form three-centre sweep (all k_three_center and
# k_herm_finish dispatches), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# runs scaled by a 256 MiB device copy's counter readings (fetch x2.00,
# write x1.00): 4.39 GB fetched + 2.12 GB written.  Counters cannot be read
# inside a timed run, so the figure is carried here; other workloads report
# null.
SWEEP_HBM_BYTES = {"si64": 6.5101e9}


def cpu_baseline(sysin, total_triples, stride, threads=1):
    """Oracle on `threads` host cores (OpenMP over atom pairs) over every
    stride-th atom pair, scaled by the three-centre evaluation ratio;
    LAPACK solve for all k on as many threads."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    import oracle
    o = oracle.Oracle(sysin)
    oracle.Oracle.set_pair_sample(stride, 0)
    try:
        t0 = time.perf_counter()
        S, _, _ = o.integral(oracle.OL, nthreads=threads)
        t1 = time.perf_counter()
        H, _, _ = o.integral(oracle.HAM, coef=sysin.pot_coeffs, nthreads=threads)
        t2 = time.perf_counter()
        _, _, sample_tr = o.last_counts()
    finally:
        oracle.Oracle.set_pair_sample(1, 0)
    # a full-size, well-conditioned pair of matrices for timing the solve
    n = o.n
    rng = np.random.default_rng(0)
    Z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Sm = Z @ Z.conj().T / n + np.eye(n)
    Hm = (Z + Z.conj().T) / 2
    with threadpool_limits(limits=threads):
        t3 = time.perf_counter()
        for _ in range(o.nk):
            oracle.solve(Hm, Sm)
        t4 = time.perf_counter()
    scale = total_triples / max(sample_tr, 1)
    est = ((t1 - t0) + (t2 - t1)) * scale + (t4 - t3)
    sample = (f"every {stride}th atom pair (i<=j) of the overlap and Hamiltonian sweeps "
              f"({sample_tr} of {total_triples} three-centre evaluations, "
              f"{t2 - t0:.1f} s), scaled by that ratio; plus zhegv for all {o.nk} k "
              f"at n={n} ({t4 - t3:.1f} s)")
    return est, sample


def run_form(s, precision, steps, warmup, label):
    """Time `steps` Mode-B iterations of one form after `warmup` untimed ones."""
    import numpy as np

    from olcao_amd import gpu
    g = gpu.OlcaoGPU(s, precision=precision)
    ns = g.vale_dim
    for w in range(warmup):
        g.scf_iteration(s.pot_coeffs, ns)
        print(f"[bench] {label} warmup {w + 1}/{warmup} {g.timings()}", flush=True)
    stage = []
    t0 = time.perf_counter()
    for k in range(steps):
        evals = g.scf_iteration(s.pot_coeffs, ns)
        stage.append(g.timings())
        print(f"[bench] {label} step {k + 1}/{steps} {stage[-1]}", flush=True)
    t1 = time.perf_counter()
    counts = g.work_counts()
    info = dict(sec=(t1 - t0) / steps, evals=np.asarray(evals), counts=counts,
                stage={k: float(np.mean([x[k] for x in stage])) for k in stage[0]},
                dims=(g.vale_dim, g.nk, g.n_pair_cells, g.n_prim_pairs))
    g.close()
    return info


def run_batch_form(cells, precision, steps, warmup, label, workers):
    """Per-cell isolated iterations (for the per-cell time distribution),
    then `steps` timed batched iterations of all cells after `warmup`."""
    import numpy as np

    from olcao_amd import gpu, workmodel
    ctxs = []
    t0 = time.perf_counter()
    for s in cells:
        ctxs.append(gpu.OlcaoGPU(s, precision=precision))
    print(f"[bench] {label}: {len(ctxs)} contexts set up in {time.perf_counter() - t0:.1f} s",
          flush=True)
    ns = [g.vale_dim for g in ctxs]
    single, flops, t3c = [], 0.0, 0.0
    for g, n in zip(ctxs, ns):
        g.scf_iteration(g.sysin.pot_coeffs, n)          # first call allocates
        g.scf_iteration(g.sysin.pot_coeffs, n)
        t = g.timings()
        single.append(t["total_ms"])
        flops += workmodel.sweep_flops(g.work_counts())
        t3c += t["three_center_ms"]
    print(f"[bench] {label}: isolated per-cell ms p50 {np.percentile(single, 50):.2f} "
          f"p99 {np.percentile(single, 99):.2f} sum {sum(single):.1f}", flush=True)
    for w in range(warmup):
        gpu.scf_iteration_batch(ctxs, ns, workers)
        print(f"[bench] {label} batch warmup {w + 1}/{warmup}", flush=True)
    secs, lat = [], []
    for k in range(steps):
        t0 = time.perf_counter()
        evals, ms = gpu.scf_iteration_batch(ctxs, ns, workers)
        secs.append(time.perf_counter() - t0)
        lat.append(ms)
        print(f"[bench] {label} batch step {k + 1}/{steps} {secs[-1] * 1e3:.1f} ms", flush=True)
    for g in ctxs:
        g.close()
    lat = np.concatenate(lat)
    return dict(sec=float(np.mean(secs)), evals=[np.asarray(e) for e in evals],
                single_ms=np.asarray(single), latency_ms=lat, flops=flops, t3c_ms=t3c)


def main_batch(args):
    import numpy as np

    from olcao_amd import configs
    cells = configs.uneven_batch()
    fp64 = run_batch_form(cells, 0, args.steps, args.warmup, "fp64", args.workers)
    mixed = None if args.fp64_only else run_batch_form(cells, 1, args.steps, args.warmup,
                                                       "mixed", args.workers)
    if mixed is None:
        print(json.dumps({"workload": "uneven_batch", "fp64_s_per_iter": fp64["sec"]}), flush=True)
        return
    HARTREE_EV = 27.211386245988
    eig_err, de = 0.0, 0.0
    for s, a, b in zip(cells, mixed["evals"], fp64["evals"]):
        n_el = sum(s.pot_types[t].Z for t in s.pot_type) - 2 * s.core_dim
        n_occ = min(s.vale_dim, int(round(n_el / 2)))
        a, b = a.reshape(s.nk, -1), b.reshape(s.nk, -1)
        eig_err = max(eig_err, float(np.abs(a[:, :n_occ] - b[:, :n_occ]).max()))
        de = max(de, abs(band_energy(a, s.kweights, n_occ) - band_energy(b, s.kweights, n_occ))
                 / s.n_atoms)
    pct = lambda a, q: float(np.percentile(a, q))   # noqa: E731
    achieved = mixed["flops"] / (mixed["t3c_ms"] * 1e-3) / 1e12
    out = {
        "metric": METRIC, "value": mixed["sec"], "unit": "s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": mixed["sec"] * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (f32 exponential in the three-centre items)", "data": "synthetic (seeded)",
        "config": {"workload": WORKLOADS["uneven_batch"], "cells": len(cells),
                   "form": "mixed (value) beside plain FP64", "mode": "B, all cells per step",
                   "workers": args.workers,
                   "atoms": int(sum(s.n_atoms for s in cells)),
                   "orbital_k_points": int(sum(s.vale_dim * s.nk for s in cells)),
                   "fp64_s_per_iter": fp64["sec"],
                   "mixed_speedup_vs_fp64": fp64["sec"] / mixed["sec"],
                   "cells_per_s_mixed": len(cells) / mixed["sec"],
                   "cells_per_s_fp64": len(cells) / fp64["sec"],
                   "isolated_sum_s_fp64": float(fp64["single_ms"].sum()) * 1e-3,
                   "isolated_cell_ms_fp64": {"p50": pct(fp64["single_ms"], 50),
                                             "p99": pct(fp64["single_ms"], 99)},
                   "isolated_cell_ms_mixed": {"p50": pct(mixed["single_ms"], 50),
                                              "p99": pct(mixed["single_ms"], 99)},
                   "batched_cell_latency_ms_mixed": {"p50": pct(mixed["latency_ms"], 50),
                                                     "p99": pct(mixed["latency_ms"], 99)},
                   "mixed_vs_fp64_max_eig_err_meV_occupied": eig_err * HARTREE_EV * 1e3,
                   "mixed_vs_fp64_band_energy_err_Ha_per_atom": de},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_three_center (all classes, every cell's sweep run alone, "
                               "mixed form)",
                     "flops_per_launch": mixed["flops"], "launch_ms": mixed["t3c_ms"]},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def band_energy(w, kweights, n_occ):
    import numpy as np
    return float(np.sum(np.asarray(kweights)[:, None] * np.asarray(w)[:, :n_occ]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=16)
    ap.add_argument("--workload", default="si64", choices=sorted(WORKLOADS),
                    help="si64 is the benchmark workload; the others are measured the same "
                         "way for the design notes")
    ap.add_argument("--fp64-only", action="store_true",
                    help="time the FP64 form only (for counter runs)")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time the mixed form only (for counter runs)")
    ap.add_argument("--workers", type=int, default=8,
                    help="host threads of the batched call (uneven_batch workload)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES once on the device (the reference kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is described, not run).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "uneven_batch":
        main_batch(args)
        return

    import numpy as np

    from olcao_amd import configs, gpu, workmodel

    s = getattr(configs, args.workload)()
    print(f"[bench] setting up {s.name}: valeDim={s.vale_dim} nk={s.nk}", flush=True)
    if args.mixed_only:
        mixed = run_form(s, gpu.PREC_MIXED, args.steps, args.warmup, "mixed")
        print(json.dumps({"workload": args.workload, "mixed_s_per_iter": mixed["sec"],
                          "stage_ms_mixed": mixed["stage"]}), flush=True)
        return
    fp64 = run_form(s, gpu.PREC_FP64, args.steps, args.warmup, "fp64")
    if args.fp64_only:
        print(json.dumps({"workload": args.workload, "fp64_s_per_iter": fp64["sec"],
                          "stage_ms_fp64": fp64["stage"]}), flush=True)
        return
    mixed = run_form(s, gpu.PREC_MIXED, args.steps, args.warmup, "mixed")
    HARTREE_EV = 27.211386245988
    n_el = sum(s.pot_types[t].Z for t in s.pot_type) - 2 * s.core_dim
    n_occ = min(s.vale_dim, int(round(n_el / 2)))
    eig_err_mev = float(np.abs(mixed["evals"][:, :n_occ] - fp64["evals"][:, :n_occ]).max()) \
        * HARTREE_EV * 1e3
    eig_err_all_mev = float(np.abs(mixed["evals"] - fp64["evals"]).max()) * HARTREE_EV * 1e3
    de = abs(band_energy(mixed["evals"], s.kweights, n_occ) -
             band_energy(fp64["evals"], s.kweights, n_occ)) / s.n_atoms
    counts = mixed["counts"]
    flops = workmodel.sweep_flops(counts)
    tri = int(counts.sum())
    t3c = mixed["stage"]["three_center_ms"] * 1e-3
    achieved = flops / t3c / 1e12
    vdim, nk, npc, npp = mixed["dims"]
    out = {
        "metric": METRIC,
        "value": mixed["sec"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["sec"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (f32 exponential in the three-centre items)",
        "data": "synthetic (seeded)",
        "config": {"workload": WORKLOADS[args.workload],
                   "form": "mixed (value) beside plain FP64", "mode": "B (full rebuild per iteration)",
                   "valeDim": vdim, "nk": nk, "pair_cells": npc, "primitive_pairs": npp,
                   "three_center_evals": tri,
                   "fp64_s_per_iter": fp64["sec"],
                   "mixed_speedup_vs_fp64": fp64["sec"] / mixed["sec"],
                   "mixed_vs_fp64_max_eig_err_meV_occupied": eig_err_mev,
                   "mixed_vs_fp64_max_eig_err_meV_all_states": eig_err_all_mev,
                   "occupied_states": n_occ,
                   "mixed_vs_fp64_band_energy_err_Ha_per_atom": de,
                   "stage_ms_mixed": mixed["stage"], "stage_ms_fp64": fp64["stage"]},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": SWEEP_HBM_BYTES.get(args.workload),
                     "kernel": "k_three_center (all classes, one sweep, mixed form)",
                     "flops_per_launch": flops, "launch_ms": t3c * 1e3},
    }
    if not args.no_cpu:
        est, sample = cpu_baseline(s, tri, args.cpu_stride)
        out["cpu_baseline"] = {"value": est, "unit": "s", "cores": 1, "kind": "port",
                               "sample": sample}
        # the same port on all host cores this process may use (at most 16)
        cores = min(len(os.sched_getaffinity(0)), 16)
        est, sample = cpu_baseline(s, tri, args.cpu_stride, threads=cores)
        out["cpu_baseline_all_cores"] = {"value": est, "unit": "s", "cores": cores,
                                         "kind": "port", "sample": sample}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
