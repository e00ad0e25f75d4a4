"""Writes kron_ref_vectors.npz: tuples produced by the reference's own Graph500
generator (graph500-generator/, compiled from its source files into
oracle/_ref/ by `make -C oracle ref`) at several scales and tuple offsets.

Run in a container that has the reference tree:  python tests/golden/make_kron_vectors.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

# (scale, first tuple, count): seeds (1, 2) as generator_test_mpi.c:77 passes them
CASES = [(10, 0, 4096), (16, 0, 2048), (20, 0, 2048), (21, 0, 64), (24, 123456789 % (16 << 24), 1024),
         (25, 500000000, 1024), (28, 4000000000, 1024), (30, 17000000000, 512)]


def main():
    L = oracle.ref_generator()
    if L is None:
        sys.exit("oracle/_ref/libref_graph500.so missing: run `make -C oracle ref` with the reference tree present")
    out = {}
    for s, e0, cnt in CASES:
        out[f"s{s}_e{e0}"] = oracle.ref_kron(L, s, e0, e0 + cnt)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "kron_ref_vectors.npz"), **out)
    print("wrote", len(out), "cases")


if __name__ == "__main__":
    main()
