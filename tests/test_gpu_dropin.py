"""The drop-in boundary, exercised from the reference's own host code.

1. oracle/_ref/ref_igc: the FillAdjListGPU replacement body of INTEGRATION.md,
   compiled against the reference's AdjList and linked to libigc.so, followed by
   the reference's RowsBySum and Codebook.
2. oracle/_ref/ref_driver gpu: the reference's unmodified CodebookAdjList with
   useGPU = true, whose FillAdjListGPU runs $INDEXGEN_ROOT/src/gpu_graph_generator.py
   -- here indexgen_amd/gpu_graph_generator.py -- and reads its pair file back.

Both must give, on the n=11 plain workload (14,409 strings, editDist 3), the
reference CPU path's result with one thread (FillAdjList + Codebook,
tests/golden/full_size_golden.npz "plain_T1"): the same 2|E| and the same
codebook, word for word in selection order.
"""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import indexgen_amd as igc
from indexgen_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")


@pytest.fixture(scope="module")
def plain(tmp_path_factory):
    z = np.load(os.path.join(ROOT, "tests", "golden", "full_size_golden.npz"))
    d = tmp_path_factory.mktemp("dropin")
    syms = W.plain_small(seed=1)
    cf = d / "cands.txt"
    cf.write_text("\n".join(igc.to_strings(syms)) + "\n")
    return {"cands": str(cf), "dir": d, "ones": int(z["plain_T1_ones"]),
            "codebook": z["plain_T1_codebook_sha256"].item().decode(),
            "codewords": int(z["plain_T1_codewords"])}


def _check(out_dir, stdout, plain):
    words = [w for w in open(os.path.join(out_dir, "codebook.txt")).read().split("\n") if w]
    line = [l for l in stdout.split("\n") if l.startswith("ones ")][-1]  # the script prints its own line first
    ones = int(line.split()[1])
    assert ones == plain["ones"]
    assert len(words) == plain["codewords"]
    assert hashlib.sha256("\n".join(words).encode()).hexdigest() == plain["codebook"]


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_igc")), reason="oracle/_ref/ref_igc not built")
def test_reference_host_linked_to_library(plain):
    out_dir = plain["dir"] / "linked"
    out_dir.mkdir()
    r = subprocess.run([os.path.join(REF, "ref_igc"), plain["cands"], "3", str(out_dir)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _check(out_dir, r.stdout, plain)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_driver")), reason="oracle/_ref/ref_driver not built")
def test_reference_gpu_path_through_script(plain):
    root = plain["dir"] / "indexgen_root"
    (root / "src").mkdir(parents=True)
    os.symlink(os.path.join(ROOT, "indexgen_amd", "gpu_graph_generator.py"), root / "src" / "gpu_graph_generator.py")
    out_dir = plain["dir"] / "script"
    out_dir.mkdir()
    env = dict(os.environ, INDEXGEN_ROOT=str(root), INDEXGEN_PYTHON=sys.executable)
    r = subprocess.run([os.path.join(REF, "ref_driver"), "gpu", "3", plain["cands"], str(out_dir)],
                       capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    _check(out_dir, r.stdout, plain)
    assert not [f for f in os.listdir(out_dir) if f.startswith("temp_")]  # the host removed its files
