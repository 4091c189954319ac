"""The codebook file igc_write_codebook writes, byte for byte against the file
the compiled reference's own ToFile writes (src/Utils.cpp:311-418, through
oracle/_ref/ref_driver header) for the same parameters and timings: a plain run
(Number of Ones in Matrix; fill / process times) and a clustered run, converged
(clustering lines; total and per-iteration clustering / solving times) and not
(clustering lines; fill / process times).  CPU only."""
import os
import subprocess

import numpy as np
import pytest

import indexgen_amd as igc
from oracle import oracle as O

pytestmark = pytest.mark.skipif(not os.path.exists(O.REF_DRIVER),
                                reason="oracle/_ref/ref_driver not built (make -C oracle ref)")

CASES = [
    # (n, min_ed, max_run, min_gc, max_gc, threads, cand_num, ones, K, iters, conv, method, times)
    (11, 3, 3, 0.3, 0.7, 32, 14409, 42880, -1, -1, 3, "hierarchical_kmeans", (0.12, 6.456, 0.184, 6.9)),
    (16, 5, 3, 0.25, 0.75, 16, 500000, 0, 256, 4, 3, "kmeans_fp32", (1.5, 12.345, 3.2, 17.0)),
    (13, 4, 0, 0.0, 0.0, 8, 200000, 0, 30, 0, 2, "random_cluster", (0.0, 2.005, 1.0, 3.5)),
    (10, 2, 2, 0.35, 0.65, 1, 7, 0, 5, 7, 7, "random_cluster", (0.001, 0.0149, 0.0051, 0.02)),
]


@pytest.mark.parametrize("case", CASES, ids=["plain", "clustered", "clustered_unconverged", "clustered_odd"])
def test_codebook_file_equals_reference(tmp_path, case):
    n, min_ed, max_run, min_gc, max_gc, T, cand_num, ones, K, iters, conv, method, (cs, fs, ps, alls) = case
    rng = np.random.default_rng(n * 100 + K)
    syms = rng.integers(0, 4, size=(40, n), dtype=np.uint8)
    order = rng.permutation(40)[:17].astype(np.uint32)
    words = igc.to_strings(syms[order])
    ref_dir = tmp_path / "ref"
    ours_dir = tmp_path / "ours"
    ref_dir.mkdir()
    ours_dir.mkdir()
    wf = tmp_path / "words.txt"
    wf.write_text("\n".join(words) + "\n")
    cand_file = "cands.txt"
    out = subprocess.run([O.REF_DRIVER, "header", str(wf), str(n), str(min_ed), str(max_run), repr(min_gc),
                          repr(max_gc), str(T), cand_file, str(cand_num), str(ones), str(K), str(iters), str(conv),
                          method, repr(cs), repr(fs), repr(ps), repr(alls), str(ref_dir)],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    ref_name = out.stdout.split()[-1]
    path = igc.write_codebook(str(ours_dir), syms, order, code_len=n, min_ed=min_ed, max_run=max_run, min_gc=min_gc,
                              max_gc=max_gc, method_line=f"Using Generation Method: FileRead (file={cand_file})",
                              candidate_num=cand_num, matrix_ones=ones, threads=T, candidates_s=cs, fill_s=fs,
                              process_s=ps, overall_s=alls, cluster_k=K, cluster_iterations=iters,
                              convergence_iterations=conv, cluster_method=method)
    assert os.path.basename(path) == ref_name
    ref = (ref_dir / ref_name).read_bytes()
    ours = open(path, "rb").read()
    assert ours == ref, "\n".join(l for l in __import__("difflib").unified_diff(
        ref.decode().split("\n"), ours.decode().split("\n"), "reference", "igc", lineterm=""))
    if K > 0:
        assert b"Clustering Method:\t\t" + method.encode() in ours and b"Number of Ones" not in ours
        assert (b"Total Clustering Time" in ours) == (iters > 0)
