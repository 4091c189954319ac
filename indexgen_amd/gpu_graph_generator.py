"""Drop-in for the reference's per-call GPU script.

The reference's FillAdjListGPU (src/SparseMat.cpp:454-534) runs

    $INDEXGEN_PYTHON $INDEXGEN_ROOT/src/gpu_graph_generator.py <cands.txt> <edges.bin> <minED> [maxGPUMemoryGB]

and reads <edges.bin> back with AdjList::FromBinaryFile.  Placed at that path
(a copy or a symlink, also as gpu_graph_generator_binary.py for isBinary runs:
strings longer than 32 take symbols 0/1 and are built the same way), this
script builds the graph with libigc.so and writes the same file: little-endian
int32 pairs (u, v), u < v, one per edge (src/gpu_graph_generator.py:57-77).
The reference host stays unmodified.  maxGPUMemoryGB is accepted and ignored:
the library sizes its buffers from the graph.  Exits 1 on any error, as the
reference's script does.
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.realpath(__file__))))


def main(argv) -> int:
    if len(argv) < 4:
        print("Usage: gpu_graph_generator.py <input_vectors.txt> <output_edges.bin> <threshold> [max_gpu_memory_gb]")
        return 1
    import indexgen_amd as igc
    try:
        with open(argv[1]) as f:
            cands = [line.strip() for line in f if line.strip()]
        ctx = igc.Context(0)
        adj = ctx.graph_build(igc.to_symbols(cands) if cands else [], int(argv[3]))
        igc.write_pair_file(adj, argv[2])
        print(f"[igc] {len(cands)} candidates, {adj.n_edges} edges -> {argv[2]}")
    except Exception as e:  # the reference's script exits 1 on any failure
        print(f"[igc] failed: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
