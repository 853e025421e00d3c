"""Writes the CSR golden fixtures tests/golden/ref_csr_<nx>x<ny>x<nz>.npz.

Runs oracle/_ref/ref_golden -- built by `make -C oracle ref` from the
reference's own NM_Mtx_Sparse.f90 (NM_CSR_AssembleFromElements ->
NM_COO_to_CSR -> Sort_CSR_Rows, and NM_CSR_MatVec_Optimized) -- in the build
container, and stores what it printed:

  ia, ja : 1-based CSR row pointers / column indices exactly as the reference
           assembler produced them (int32)
  a      : values for the integer element matrices Ke(i,j) = 1 + (7e+3i+j) mod 11
           (1-based e, i, j), exact in FP64
  y      : NM_CSR_MatVec_Optimized(A, x) with x(i) = 1 + (5i) mod 13 (1-based i)
  yt     : A^T x with A^T from NM_CSR_Transpose (NM_Mtx_SpMVCSRTranspose.f90:46-145)

Usage:  python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_golden")
MESHES = [(1, 1, 1), (2, 1, 1), (3, 2, 2), (20, 4, 4)]


def run(nx, ny, nz):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "g.txt")
        subprocess.run([EXE, str(nx), str(ny), str(nz), out], check=True)
        with open(out) as f:
            n, nnz = (int(t) for t in f.readline().split())
            ia = np.array([int(f.readline()) for _ in range(n + 1)], np.int32)
            ja = np.empty(nnz, np.int32)
            a = np.empty(nnz, np.float64)
            for k in range(nnz):
                s = f.readline().split()
                ja[k] = int(s[0])
                a[k] = float(s[1])
            y = np.array([float(f.readline()) for _ in range(n)], np.float64)
            yt = np.array([float(f.readline()) for _ in range(n)], np.float64)
    np.savez_compressed(os.path.join(HERE, f"ref_csr_{nx}x{ny}x{nz}.npz"),
                        ia=ia, ja=ja, a=a, y=y, yt=yt, mesh=np.array([nx, ny, nz], np.int32))
    return n, nnz


if __name__ == "__main__":
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    for m in MESHES:
        print(m, run(*m), file=sys.stderr)
