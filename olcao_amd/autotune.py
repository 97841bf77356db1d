"""Autotuner for the tile shapes of the integral build.

Three shapes decide how the build's work is cut into GPU work units, and
their best values depend on the system:

  chunk_items  three-centre items per image-list chunk (one lane of the
               three-centre kernels each): small chunks balance diffuse pairs
               with thousands of images against compact ones, large chunks
               amortise the per-chunk Hermite-sum traffic
  seg_pairs    primitive pairs per contraction wavefront (1..64): short
               segments spread on-site pair-cells (hundreds of pairs) over
               more wavefronts, long ones need fewer partial sums
  bloch        the Bloch-sum kernel: the matrix-core one (FP64 MFMA, or the
               split BF16 products in the mixed form), the FP64 MFMA one in
               either form, or the vector-ALU one

The tuner measures the build (every stage of a Mode-B iteration before the
eigensolve, HIP events on the context stream) for each candidate, one shape
at a time from the defaults (coordinate descent, two passes), and keeps the
fastest.  Results go into a JSON table keyed by a system signature and the
precision; `lookup` returns the entry of the nearest signature, so a system
the table does not hold takes the shapes tuned for the most similar one.

    python -m olcao_amd.autotune --workload si64 quartz diamond_c8 \\
        --out olcao_amd/tuning.json
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

CANDIDATES = {
    "chunk_items": [64, 128, 192, 384, 768],
    "seg_pairs": [8, 16, 32, 64],
    # OLCAO_BLOCH_VALU, OLCAO_BLOCH_MFMA, OLCAO_BLOCH_MFMA64 (the last the
    # same kernel as the second in the FP64 form, so tried only for mixed)
    "bloch": [1, 2, 3],
}
DEFAULTS = {"chunk_items": 192, "seg_pairs": 32, "bloch": 2}


def signature(sysin) -> dict:
    """What the shapes depend on: orbitals, k-points, atoms, real or complex."""
    return {"n": int(sysin.vale_dim), "nk": int(sysin.nk), "atoms": int(sysin.n_atoms),
            "gamma": bool(sysin.gamma)}


def _key(sig: dict, precision: int) -> str:
    return (f"n{sig['n']}-k{sig['nk']}-a{sig['atoms']}-{'real' if sig['gamma'] else 'cplx'}"
            f"-p{precision}")


def _distance(a: dict, b: dict) -> float:
    return (abs(math.log(a["n"] / b["n"])) + abs(math.log(a["nk"] / b["nk"]))
            + abs(math.log(a["atoms"] / b["atoms"])) + (0.0 if a["gamma"] == b["gamma"] else 1.0))


def load(path: str = TABLE) -> dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def lookup(sysin, precision: int, path: str = TABLE):
    """Tuned shapes for this system and precision: the table entry of the
    same precision whose signature is nearest (None when the table holds
    none)."""
    table = load(path)
    sig = signature(sysin)
    best, dist = None, None
    for e in table.values():
        if e.get("precision") != precision:
            continue
        d = _distance(sig, e["signature"])
        if dist is None or d < dist:
            best, dist = e, d
    return None if best is None else dict(best["shapes"])


def _build_ms(sysin, precision, shapes, reps):
    """Median build time (ms) of `reps` Mode-B iterations after one warm-up,
    with the given shapes."""
    import numpy as np

    from olcao_amd import gpu
    g = gpu.OlcaoGPU(sysin, precision=precision, chunk_items=shapes["chunk_items"],
                     seg_pairs=shapes["seg_pairs"], bloch=shapes["bloch"])
    try:
        n = g.vale_dim
        g.scf_iteration(sysin.pot_coeffs, n)
        t = []
        for _ in range(reps):
            g.scf_iteration(sysin.pot_coeffs, n)
            d = g.timings()
            t.append(d["total_ms"] - d["solve_ms"])
        return float(np.median(t))
    finally:
        g.close()


def tune(sysin, precision: int, reps: int = 5, passes: int = 2, log=print) -> dict:
    """Coordinate descent over CANDIDATES from DEFAULTS; returns the table
    entry (signature, precision, shapes, build_ms, default_ms, trials)."""
    shapes = dict(DEFAULTS)
    trials = []
    cache = {}

    def measure(s):
        k = (s["chunk_items"], s["seg_pairs"], s["bloch"])
        if k not in cache:
            t0 = time.perf_counter()
            cache[k] = _build_ms(sysin, precision, s, reps)
            trials.append({**s, "build_ms": cache[k]})
            log(f"[tune] {sysin.name} p{precision} {s}: build {cache[k]:.3f} ms "
                f"({time.perf_counter() - t0:.1f} s)")
        return cache[k]

    default_ms = measure(shapes)
    best = default_ms
    for _ in range(passes):
        changed = False
        for name, values in CANDIDATES.items():
            for v in values:
                if v == shapes[name] or (name == "bloch" and v == 3 and precision == 0):
                    continue
                cand = dict(shapes, **{name: v})
                t = measure(cand)
                if t < best * 0.99:          # keep a change only if it gains >= 1%
                    best, shapes, changed = t, cand, True
        if not changed:
            break
    return {"signature": signature(sysin), "precision": precision, "shapes": shapes,
            "build_ms": best, "default_ms": default_ms, "trials": trials}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", nargs="+", default=["si64"],
                    help="configs.<name>() systems; 'uneven_batch:i,j' takes cells i, j of "
                         "the uneven batch")
    ap.add_argument("--precision", nargs="+", type=int, default=[0, 1])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=TABLE)
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    from olcao_amd import configs
    systems = []
    for w in args.workload:
        if w.startswith("uneven_batch:"):
            cells = configs.uneven_batch()
            systems += [cells[int(i)] for i in w.split(":", 1)[1].split(",")]
        else:
            systems.append(getattr(configs, w)())
    table = load(args.out)
    for s in systems:
        for p in args.precision:
            e = tune(s, p, args.reps)
            table[_key(e["signature"], p)] = e
            print(f"[tune] {s.name} p{p}: {e['shapes']} build {e['build_ms']:.3f} ms "
                  f"(defaults {e['default_ms']:.3f} ms)", flush=True)
            with open(args.out, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
