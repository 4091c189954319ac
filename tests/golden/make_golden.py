"""Writes the golden fixtures of tests/golden from the compiled IndexGen reference.

Run in the build container (where /root/reference exists) after
`make -C oracle ref`:

    python tests/golden/make_golden.py

Outputs (data only -- inputs and the reference's outputs):
  pairs_golden.npz      random / near / exhaustive-short string pairs with the
                        reference's EditDistanceBanded(text, MakePattern(pat), k)
                        and FastEditDistance values
  graphs_golden.npz     small candidate sets with the reference's FillAdjList
                        adjacency (AdjList::ToFile list order) and Codebook
                        selection order, for T = 1 and T = 4 threads
  published_*.npz       candidates of three codebooks the reference publishes
                        (Codebooks/N11/...), regenerated with the reference's
                        LinearCode generator, with the published edge count
                        ("Number of Ones in Matrix") and codebook digest
"""
from __future__ import annotations

import hashlib
import itertools
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

PUBLISHED = [
    "Codebooks/N11/N11-ED5-GC0.3-0.7-RUN0.txt",
    "Codebooks/N11/N11-ED4-GC0.3-0.7-RUN0.txt",
    "Codebooks/N11/N11-ED3-GC0-0-RUN0.txt",
]


def sym_str(row) -> str:
    return "".join(chr(48 + int(c)) for c in row)


def make_pairs(rng: np.random.Generator):
    lines = []
    # exhaustive for n <= 3 and every k, equal and unequal lengths 0..3
    alphabet = "0123"
    for la in range(1, 4):
        for lb in range(1, 4):
            for a in itertools.product(alphabet, repeat=la):
                for b in itertools.product(alphabet, repeat=lb):
                    for k in range(0, 5):
                        lines.append((k, "".join(a), "".join(b)))
    # random and near pairs, equal length 4..16
    for n in range(4, 17):
        for t in range(1500):
            a = rng.integers(0, 4, n)
            if t % 2:
                b = rng.integers(0, 4, n)
            else:
                b = a.copy()
                for _ in range(int(rng.integers(1, 6))):
                    op = int(rng.integers(0, 3))
                    p = int(rng.integers(0, n))
                    if op == 0:
                        b[p] = rng.integers(0, 4)
                    elif op == 1:
                        b = np.concatenate([np.delete(b, p), rng.integers(0, 4, 1)])
                    else:
                        b = np.concatenate([b[:p], rng.integers(0, 4, 1), b[p:]])[:n]
            lines.append((int(rng.integers(0, 8)), sym_str(b), sym_str(a)))
    inp = "".join(f"{k} {t} {p}\n" for k, t, p in lines)
    out = subprocess.run([DRIVER, "pairs"], input=inp, capture_output=True, text=True, check=True).stdout.split()
    res = np.array(out, dtype=np.int32).reshape(-1, 2)
    L = 16
    text = np.full((len(lines), L), 255, np.uint8)
    pat = np.full((len(lines), L), 255, np.uint8)
    tl = np.zeros(len(lines), np.int32)
    pl = np.zeros(len(lines), np.int32)
    for r, (k, t, p) in enumerate(lines):
        text[r, :len(t)] = [int(c) for c in t]
        pat[r, :len(p)] = [int(c) for c in p]
        tl[r], pl[r] = len(t), len(p)
    ks = np.array([k for k, _, _ in lines], np.int32)
    np.savez_compressed(os.path.join(HERE, "pairs_golden.npz"), k=ks, text=text, text_len=tl, pat=pat, pat_len=pl,
                        banded=res[:, 0], wf=res[:, 1])
    print("pairs:", len(lines))


def ref_graph(syms: np.ndarray, min_ed: int, T: int):
    d = tempfile.mkdtemp(prefix="igc_golden_")
    cf = os.path.join(d, "c.txt")
    with open(cf, "w") as f:
        for r in syms:
            f.write(sym_str(r) + "\n")
    out = subprocess.run([DRIVER, "graph", str(T), str(min_ed), cf, d], capture_output=True, text=True,
                         check=True).stdout.split()
    ones = int(out[1])
    adj = np.loadtxt(os.path.join(d, "adj.txt"), dtype=np.int64, ndmin=2).reshape(-1, 2)
    N = syms.shape[0]
    deg = np.bincount(adj[:, 0], minlength=N) if adj.size else np.zeros(N, np.int64)
    row_ptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = adj[:, 1].astype(np.uint32)
    words = [w.strip() for w in open(os.path.join(d, "codebook.txt")) if w.strip()]
    # codewords as symbol rows (indices would be ambiguous for repeated strings)
    order = np.array([[int(c) for c in w] for w in words], np.uint8).reshape(len(words), syms.shape[1])
    for fn in os.listdir(d):
        os.remove(os.path.join(d, fn))
    os.rmdir(d)
    assert ones == col.size
    return row_ptr, col, order


def make_graphs(rng: np.random.Generator):
    specs = [(300, 6, 3), (800, 10, 4), (1200, 11, 3), (600, 13, 5), (500, 16, 5), (400, 8, 2), (350, 9, 6),
             (200, 5, 1)]
    arrays = {}
    for s, (N, n, ed) in enumerate(specs):
        syms = rng.integers(0, 4, size=(N, n), dtype=np.uint8)
        if s == 7:
            syms[50:60] = syms[0]  # repeated strings: distance 0 edges
        else:
            _, idx = np.unique(syms, axis=0, return_index=True)
            syms = syms[np.sort(idx)]
        arrays[f"s{s}_syms"] = syms
        arrays[f"s{s}_min_ed"] = np.int32(ed)
        for T in (1, 4):
            rp, col, order = ref_graph(syms, ed, T)
            arrays[f"s{s}_T{T}_row_ptr"] = rp
            arrays[f"s{s}_T{T}_col"] = col
            arrays[f"s{s}_T{T}_codebook"] = order
        print("graph set", s, syms.shape, "edges", arrays[f"s{s}_T1_col"].size // 2)
    arrays["count"] = np.int32(len(specs))
    np.savez_compressed(os.path.join(HERE, "graphs_golden.npz"), **arrays)


def parse_published(path: str):
    txt = open(path).read()
    head, body = txt.split("===========================================", 1)
    g = lambda pat: re.search(pat, head).group(1)
    info = {
        "n": int(g(r"Code Length:\s+(\d+)")),
        "min_ed": int(g(r"Min Codebook Edit Distance:\s+(\d+)")),
        "max_run": int(g(r"Max Run:\s+(\d+)")),
        "min_gc": float(g(r"Min GC Content:\s+([\d.]+)")),
        "max_gc": float(g(r"Max GC Content:\s+([\d.]+)")),
        "min_hd": int(g(r"LinearCode \(minHD=(\d+)\)")),
        "bias": g(r"Bias vector: \[([\d, ]*)\]").replace(" ", ""),
        "row_perm": g(r"Row permutation: \[([\d, ]*)\]").replace(" ", ""),
        "col_perm": g(r"Column permutation: \[([\d, ]*)\]").replace(" ", ""),
        "cands": int(g(r"Number of Candidates:\s+(\d+)")),
        "ones": int(g(r"Number of Ones in Matrix:\s+(\d+)")),
        "size": int(g(r"Number of Code Words:\s+(\d+)")),
        "threads": int(g(r"Number of Threads:\s+(\d+)")),
    }
    words = [w for w in body.split() if w]
    assert len(words) == info["size"]
    return info, words


def make_published():
    for rel in PUBLISHED:
        info, words = parse_published(os.path.join(REF, rel))
        out = subprocess.run([DRIVER, "linear", str(info["n"]), str(info["min_hd"]), info["bias"] or "-",
                              info["row_perm"] or "-", info["col_perm"] or "-", str(info["max_run"]),
                              str(info["min_gc"]), str(info["max_gc"])], capture_output=True, text=True,
                             check=True).stdout.split()
        assert len(out) == info["cands"], (rel, len(out), info["cands"])
        syms = np.array([[int(c) for c in w] for w in out], np.uint8)
        index = {w: i for i, w in enumerate(out)}
        order = np.array([index[w] for w in words], np.int32)
        digest = hashlib.sha256("\n".join(words).encode()).hexdigest()
        name = os.path.basename(rel)[:-4].replace(".", "_")
        np.savez_compressed(os.path.join(HERE, f"published_{name}.npz"), syms=syms, min_ed=np.int32(info["min_ed"]),
                            ones=np.int64(info["ones"]), threads=np.int32(info["threads"]), order=order,
                            digest=np.bytes_(digest.encode()), source=np.bytes_(rel.encode()))
        print("published", rel, syms.shape, "ones", info["ones"], "codewords", info["size"])


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle ref")
    rng = np.random.default_rng(20251203)
    make_pairs(rng)
    make_graphs(rng)
    make_published()


if __name__ == "__main__":
    main()
