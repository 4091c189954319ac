"""Measures the two pair predicates on the GPU and prints the time tables that
pred_kind() in csrc/igc_lv.hpp chooses by.

    python -m indexgen_amd.predicate_sweep [--N 65536] [--reps 1] [--json out.json]

For each word class (two pairs per register for n <= 16, one pair per 32-bit
lane for n <= 32, binary words on 64-bit lanes for n <= 64) it forces
Landau-Vishkin at every instantiated K and Myers at several lengths, times the
count-only pair kernel at every tile shape (igc_autotune, best shape kept) on
N random strings, and reports femtoseconds per pair: Landau-Vishkin per K, and
a least-squares line a + b*n for Myers.
"""
import argparse
import json

import numpy as np

import indexgen_amd as igc

CLASSES = (
    ("u16x2", 16, (4, 8, 12, 16)),
    ("u32", 32, (17, 20, 24, 28, 32)),
    ("bin64", 64, (33, 40, 48, 56, 64)),
)
LV_MAX_K = 13


def _best_fs(ctx, n, ed, N, reps):
    ents = ctx.autotune(n, ed, N, reps=reps)
    best = max(ents, key=lambda e: e["gpairs_s"])
    return 1e6 / best["gpairs_s"], best


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--N", type=int, default=65536)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    ctx = igc.Context(0)
    out = {"N": a.N, "lv_fs": {}, "myers_fs": {}, "myers_fit": {}}
    try:
        for name, n_lv, myers_ns in CLASSES:
            ctx.set_predicate(igc.IGC_PRED_LANDAU_VISHKIN)
            row = []
            for K in range(LV_MAX_K + 1):
                fs, b = _best_fs(ctx, n_lv, K + 1, a.N, a.reps)
                row.append(fs)
                print(f"LV    {name:6s} n={n_lv:2d} K={K:2d} {fs:9.1f} fs/pair  (cpl={b['cpl']} q={b['q']})", flush=True)
            out["lv_fs"][name] = row
            ctx.set_predicate(igc.IGC_PRED_MYERS)
            pts = []
            for n in myers_ns:
                fs, b = _best_fs(ctx, n, 4, a.N, a.reps)
                pts.append((n, fs))
                print(f"MYERS {name:6s} n={n:2d}      {fs:9.1f} fs/pair  (cpl={b['cpl']} q={b['q']})", flush=True)
            x = np.array([p[0] for p in pts], float)
            y = np.array([p[1] for p in pts], float)
            slope, icpt = np.polyfit(x, y, 1)
            out["myers_fs"][name] = pts
            out["myers_fit"][name] = (icpt, slope)
    finally:
        ctx.set_predicate(igc.IGC_PRED_AUTO)
    print("lv_fs table:")
    for name, _, _ in CLASSES:
        print("    {" + ", ".join(str(int(round(v))) for v in out["lv_fs"][name]) + "},  // " + name)
    print("myers_fs a = {" + ", ".join(str(int(round(out["myers_fit"][c][0]))) for c, _, _ in CLASSES) + "}, b = {" +
          ", ".join(str(int(round(out["myers_fit"][c][1]))) for c, _, _ in CLASSES) + "}")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
