"""Benchmark: seconds per SCF iteration (H/S build + eigensolve) on one GPU.

One step is one Mode-B SCF iteration of the OLCAO path (intgPSCF + bandPSCF
semantics): the overlap sweep S(k), the fused Hamiltonian sweep H(k) with
the potential coefficients folded in (kinetic + potential terms + nuclear,
three-centre lattice sums), contraction, Bloch sums, and the batched FP64
generalised eigensolve for every k-point.  Inputs are resident on the
device before the timed region (the system and screened work lists are
built once, as the reference's setup does).

Two forms run in the same process:
  plain   the FP64 drop-in with the default tile shapes, each iteration
          launched kernel by kernel;
  chosen  the mixed form (split BF16 matrix-core Bloch sums, FP32
          exponentials of distant images) with the autotuned tile shapes
          (olcao_amd/tuning.json), each iteration replayed from a captured
          graph.
`value` is the chosen form; the plain form's time and the chosen form's
eigenvalue and band-energy errors against it are reported beside it.

Workload: si64 -- Si64 2x2x2 diamond supercell (a = 5.431 A, 0.05 A seeded
jitter), s+p+d per Si from 24 even-tempered exponents 0.05..1e7, 16
potential Gaussians 0.1..1e5 per site, 2x2x2 Monkhorst-Pack mesh (8 k),
576 orbitals; synthetic seeded data.

The JSON line carries the roofline of the largest kernel of the iteration,
the tridiagonal reduction k_trd ((16/3) n^3 real flops per complex matrix,
timed by HIP events around its launch on the context stream), the
three-centre sweep's beside it (FP64 operations of the kernels' own
McMurchie-Davidson formulation, olcao_amd.workmodel, over the device's
evaluation counters, divided by the sweep's HIP-event time), and a CPU
baseline: the oracle (a C restatement of the reference algorithm) on one
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
# HBM bytes of the tridiagonal reduction of one Si64 iteration (one launch,
# 8 matrices), measured the same way: 0.079 GB fetched + 0.253 GB written.
TRD_HBM_BYTES = {"si64": 3.321e8}


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


def _form_args(chosen):
    from olcao_amd import gpu
    return dict(precision=gpu.PREC_MIXED if chosen else gpu.PREC_FP64,
                tuning="table" if chosen else None)


def run_form(s, steps, warmup, label, chosen):
    """Time `steps` Mode-B iterations of one form after `warmup` untimed
    ones (the chosen form as graph replays).  Nothing but the iterations
    runs in the timed region."""
    import numpy as np

    from olcao_amd import gpu
    g = gpu.OlcaoGPU(s, graph=chosen, **_form_args(chosen))
    ns = g.vale_dim
    for w in range(warmup):
        g.scf_iteration(s.pot_coeffs, ns)
        print(f"[bench] {label} warmup {w + 1}/{warmup}", flush=True)
    gstate = g.graph_state()
    t0 = time.perf_counter()
    for k in range(steps):
        evals = g.scf_iteration(s.pot_coeffs, ns)
    t1 = time.perf_counter()
    last = g.timings()
    print(f"[bench] {label}: {steps} steps, {(t1 - t0) / steps * 1e3:.3f} ms per step, graph "
          f"state {gstate[0]} {gstate[1]}, last step {last['total_ms']:.3f} ms on the device",
          flush=True)
    info = dict(sec=(t1 - t0) / steps, evals=np.asarray(evals), tuning=g.tuning,
                graph=bool(last["graph_replay"]),
                dims=(g.vale_dim, g.nk, g.n_pair_cells, g.n_prim_pairs))
    g.close()
    return info


def time_stages(s, steps, chosen):
    """Stage and kernel times (HIP events on the context stream) of one form
    over `steps` eager iterations after one: the mean of each."""
    import numpy as np

    from olcao_amd import gpu
    g = gpu.OlcaoGPU(s, **_form_args(chosen))
    g.scf_iteration(s.pot_coeffs, g.vale_dim)
    st = []
    for _ in range(steps):
        g.scf_iteration(s.pot_coeffs, g.vale_dim)
        st.append(g.timings())
    out = {k: float(np.mean([x[k] for x in st])) for k in st[0] if k != "graph_replay"}
    out["counts"] = g.work_counts()
    out["class_pairs"] = g.class_pairs()
    g.close()
    return out


def run_batch_form(cells, precision, steps, warmup, label, workers, modes):
    """Per-cell isolated iterations (for the per-cell time distribution),
    then, for each batch build mode in `modes`, `steps` timed batched
    iterations of all cells after `warmup`; the form's time is the faster
    mode's."""
    import numpy as np

    from olcao_amd import gpu, workmodel
    ctxs = []
    t0 = time.perf_counter()
    for s in cells:
        ctxs.append(gpu.OlcaoGPU(s, precision=precision,
                                 tuning="table" if precision == gpu.PREC_MIXED else None))
    print(f"[bench] {label}: {len(ctxs)} contexts set up in {time.perf_counter() - t0:.1f} s",
          flush=True)
    ns = [g.vale_dim for g in ctxs]
    single, flops, t3c = [], 0.0, 0.0
    for g, n in zip(ctxs, ns):
        g.scf_iteration(g.sysin.pot_coeffs, n)          # first call allocates
        g.scf_iteration(g.sysin.pot_coeffs, n)
        t = g.timings()
        single.append(t["total_ms"])
        flops += workmodel.sweep_flops(g.work_counts(), g.class_pairs())
        t3c += t["three_center_ms"]
    print(f"[bench] {label}: isolated per-cell ms p50 {np.percentile(single, 50):.2f} "
          f"p99 {np.percentile(single, 99):.2f} sum {sum(single):.1f}", flush=True)
    res = dict(single_ms=np.asarray(single), flops=flops, t3c_ms=t3c, modes={})
    for mode in modes:
        name = "cells" if mode == gpu.BATCH_CELLS else "shared"
        for w in range(warmup):
            gpu.scf_iteration_batch(ctxs, ns, workers, mode)
            print(f"[bench] {label} batch ({name}) warmup {w + 1}/{warmup}", flush=True)
        secs, lat = [], []
        for k in range(steps):
            t0 = time.perf_counter()
            evals, ms = gpu.scf_iteration_batch(ctxs, ns, workers, mode)
            secs.append(time.perf_counter() - t0)
            lat.append(ms)
            print(f"[bench] {label} batch ({name}) step {k + 1}/{steps} {secs[-1] * 1e3:.1f} ms",
                  flush=True)
        res["modes"][name] = dict(sec=float(np.mean(secs)), evals=[np.asarray(e) for e in evals],
                                  latency_ms=np.concatenate(lat))
    for g in ctxs:
        g.close()
    # the faster build mode is the form's time
    best = min(res["modes"], key=lambda m: res["modes"][m]["sec"])
    res.update(res["modes"][best], mode=best)
    return res


def main_batch(args):
    import numpy as np

    from olcao_amd import configs
    cells = configs.uneven_batch()
    from olcao_amd import gpu
    # plain: every cell's build on its own streams; chosen: the faster of
    # that and the shared three-centre launches
    fp64 = run_batch_form(cells, 0, args.steps, args.warmup, "fp64", args.workers,
                          [gpu.BATCH_CELLS])
    mixed = None if args.fp64_only else run_batch_form(
        cells, 1, args.steps, args.warmup, "mixed", args.workers,
        [gpu.BATCH_CELLS, gpu.BATCH_SHARED])
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
        "dtype": "f64 (bf16-split Bloch sums with f32 accumulation, f32 exponentials of "
                 "distant images)", "data": "synthetic (seeded)",
        "config": {"workload": WORKLOADS["uneven_batch"], "cells": len(cells),
                   "form": "mixed (value) beside plain FP64", "mode": "B, all cells per step",
                   "workers": args.workers,
                   "atoms": int(sum(s.n_atoms for s in cells)),
                   "orbital_k_points": int(sum(s.vale_dim * s.nk for s in cells)),
                   "fp64_s_per_iter": fp64["sec"],
                   "mixed_batch_mode": mixed["mode"],
                   "mixed_s_per_iter_by_batch_mode": {m: v["sec"] for m, v in mixed["modes"].items()},
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
    if args.mixed_only or args.fp64_only:   # counter runs: one form, eager
        st = time_stages(s, args.steps, args.mixed_only)
        print(json.dumps({"workload": args.workload, "form": "chosen" if args.mixed_only else "plain",
                          "stage_ms": {k: v for k, v in st.items() if k.endswith("_ms")}}),
              flush=True)
        return
    plain = run_form(s, args.steps, args.warmup, "plain", False)
    chosen = run_form(s, args.steps, args.warmup, "chosen", True)
    st_plain = time_stages(s, 3, False)
    st_chosen = time_stages(s, 3, True)
    HARTREE_EV = 27.211386245988
    n_el = sum(s.pot_types[t].Z for t in s.pot_type) - 2 * s.core_dim
    n_occ = min(s.vale_dim, int(round(n_el / 2)))
    ev_c = chosen["evals"]
    eig_err_mev = float(np.abs(ev_c[..., :n_occ] - plain["evals"][..., :n_occ]).max()) \
        * HARTREE_EV * 1e3
    eig_err_all_mev = float(np.abs(ev_c - plain["evals"]).max()) * HARTREE_EV * 1e3
    de = abs(band_energy(ev_c, s.kweights, n_occ) -
             band_energy(plain["evals"], s.kweights, n_occ)) / s.n_atoms
    vdim, nk, npc, npp = plain["dims"]
    # roofline of the largest kernel: the tridiagonal reduction
    cplx = not s.gamma
    trd_ms = st_plain["trd_ms"]
    trd_flops = workmodel.hetrd_flops(vdim, nk * s.spin, cplx)
    # the three-centre sweep, by the kernels' own formulation
    counts = st_plain["counts"]
    sweep_flops = workmodel.sweep_flops(counts, st_plain["class_pairs"])
    t3c = st_plain["three_center_ms"]
    tri = int(np.asarray(counts).sum())
    if trd_ms > 0:
        roof = {"bound": "mfma", "achieved": trd_flops / (trd_ms * 1e-3) / 1e12,
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": trd_flops / (trd_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                "traffic": TRD_HBM_BYTES.get(args.workload),
                "kernel": f"k_trd (tridiagonal reduction, {nk * s.spin} x n={vdim} "
                          f"{'complex' if cplx else 'real'}; FP64 vector work, whose peak equals "
                          "the FP64 matrix peak)",
                "flops_per_launch": trd_flops, "launch_ms": trd_ms}
    else:
        roof = None
    roof_3c = {"bound": "mfma", "achieved": sweep_flops / (t3c * 1e-3) / 1e12,
               "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
               "frac": sweep_flops / (t3c * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
               "traffic": SWEEP_HBM_BYTES.get(args.workload),
               "kernel": "k_three_center + k_herm_finish (all classes, one Hamiltonian sweep, "
                         "plain form; McMurchie-Davidson operation counts)",
               "flops_per_launch": sweep_flops, "launch_ms": t3c,
               "special_evals": workmodel.sweep_specials(counts)}
    out = {
        "metric": METRIC,
        "value": chosen["sec"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": chosen["sec"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (bf16-split Bloch sums with f32 accumulation, f32 exponentials of "
                 "distant images)",
        "data": "synthetic (seeded)",
        "config": {"workload": WORKLOADS[args.workload],
                   "form": "chosen (mixed, autotuned, graph replay) beside plain FP64",
                   "mode": "B (full rebuild per iteration)",
                   "valeDim": vdim, "nk": nk, "pair_cells": npc, "primitive_pairs": npp,
                   "three_center_evals": tri,
                   "plain_fp64_s_per_iter": plain["sec"],
                   "chosen_speedup_vs_plain": plain["sec"] / chosen["sec"],
                   "chosen_tuning": chosen["tuning"], "chosen_graph_replay": chosen["graph"],
                   "chosen_vs_plain_max_eig_err_meV_occupied": eig_err_mev,
                   "chosen_vs_plain_max_eig_err_meV_all_states": eig_err_all_mev,
                   "occupied_states": n_occ,
                   "chosen_vs_plain_band_energy_err_Ha_per_atom": de,
                   "stage_ms_plain": {k: v for k, v in st_plain.items() if k.endswith("_ms")},
                   "stage_ms_chosen_eager": {k: v for k, v in st_chosen.items()
                                             if k.endswith("_ms")}},
        "roofline": roof if roof is not None else roof_3c,
        "roofline_three_center": roof_3c,
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
