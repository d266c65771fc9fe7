"""Writes the golden fixtures under tests/golden/ (run from the repo root:
python tests/golden/make_golden.py).  Committed together with its output.

  sample_4x4.npz      the reference's 4x4 sample CSR (include/loops/util/
                      sample.hxx:29-49, 65-82) with x = [1, 2, 3, 4] and the
                      closed-form y = [0, 5x0 + 8x1, 3x2, 6x1] read off the
                      logical matrix in its comment.
  chesapeake_degrees.npy
                      y for x = 1 on datasets/chesapeake/chesapeake.mtx (a
                      copy is tests/golden/chesapeake.mtx): the degree of
                      every vertex after symmetric expansion, counted
                      directly from the file's entry lines (independent of
                      any CSR conversion).  README.md:82-88 states the
                      matrix is 39 x 39 with 340 nonzeros.
  c1_4096.npz         the 4096 x 4096, 8 nnz/row synthetic matrix (seed 1):
                      CRC32 of its arrays, x (seed 11), the oracle's y, and
                      the merge-path tile coordinates at (128 threads, 5
                      items) and (256, 8).  A regression fixture for the
                      generator and the oracle.
"""
import os
import sys
import zlib

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
HERE = os.path.join(ROOT, "tests", "golden")


def sample():
    offsets = np.array([0, 0, 2, 3, 4], np.int32)
    indices = np.array([0, 1, 2, 1], np.int32)
    values = np.array([5, 8, 3, 6], np.float32)
    x = np.array([1, 2, 3, 4], np.float32)
    y = np.array([0, 5 * 1 + 8 * 2, 3 * 3, 6 * 2], np.float32)
    np.savez(os.path.join(HERE, "sample_4x4.npz"), offsets=offsets,
             indices=indices, values=values, x=x, y=y)


def chesapeake_degrees():
    deg = None
    with open(os.path.join(HERE, "chesapeake.mtx")) as f:
        lines = [l for l in f if not l.startswith("%")]
    m, n, e = map(int, lines[0].split())
    deg = np.zeros(m, np.float32)
    for l in lines[1:1 + e]:
        r, c = map(int, l.split()[:2])
        deg[r - 1] += 1
        if r != c:
            deg[c - 1] += 1
    assert m == 39 and n == 39 and int(deg.sum()) == 340
    np.save(os.path.join(HERE, "chesapeake_degrees.npy"), deg)


def c1():
    import oracle
    from edge_loops_amd import synth
    A = synth.c1_uniform_4k()
    x = synth.uniform(A.cols, seed=11)
    y = oracle.spmv_f32(A.offsets, A.indices, A.values, x)
    crc = np.array([zlib.crc32(A.offsets.tobytes()), zlib.crc32(A.indices.tobytes()),
                    zlib.crc32(A.values.tobytes()), zlib.crc32(x.tobytes())], np.uint64)
    np.savez(os.path.join(HERE, "c1_4096.npz"), crc=crc, x=x, y=y,
             coords_128_5=oracle.mp_tile_coords(A.offsets, 128 * 5),
             coords_256_8=oracle.mp_tile_coords(A.offsets, 256 * 8))


if __name__ == "__main__":
    sample()
    chesapeake_degrees()
    c1()
    print("golden fixtures written to", HERE)
