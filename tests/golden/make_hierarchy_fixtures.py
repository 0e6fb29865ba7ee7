"""Writes the per-level operators of the TEST_ij fsai job's BoomerAMG hierarchy
(ij -n 10 10 10 -solver 1 -rlx 7: 3D 7-pt Laplacian 10^3, HMIS, extended+i
P max 4, theta 0.25, max row sum 1, coarse size 9) in the layout HYPRE's
setup writes with DEBUG_SAVE_ALL_OPS (parcsr_ls/par_amg_setup.c:3118-3128):
A_00.IJ.out.00000 for the finest operator, A_%02d for the coarse ones and
P_%02d for the interpolation of each level, each written as
hypre_ParCSRMatrixPrintIJ does on one process with base 0
(parcsr_mv/par_csr_matrix.c:969-1000): the line "ilower iupper jlower
jupper", then "I J value" per stored entry in row order with %.14e values.
The hierarchy is the oracle's restatement of HYPRE's host setup (RAP in
HYPRE's host order), which reproduces the job's fixture (TEST_ij/fsai.saved:
22 iterations, final relative residual 7.480945e-09).  Files are gzipped.

Run from the repository root: python tests/golden/make_hierarchy_fixtures.py
"""
import gzip
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "hierarchy_ij_fsai103")


def print_ij(path, nrows, ncols, I, J, V):
    with gzip.open(path + ".gz", "wt") as fh:
        fh.write(f"0 {nrows - 1} 0 {ncols - 1}\n")
        for r in range(nrows):
            for p in range(I[r], I[r + 1]):
                fh.write(f"{r} {J[p]} {V[p]:.14e}\n")


def main():
    I, J, A = O.stencil(7, 10, 10, 10)
    levels = O.amg_hierarchy(1000, I, J, A, theta=0.25, max_row_sum=1.0, max_elmts=4,
                             max_coarse_size=9, coarsen="hmis", rap="kt")
    os.makedirs(OUT, exist_ok=True)
    for lvl, L in enumerate(levels):
        print_ij(os.path.join(OUT, f"A_{lvl:02d}.IJ.out.00000"), L["n"], L["n"], L["I"], L["J"],
                 L["A"])
        if "P" in L:
            PI, PJ, PA = L["P"]
            print_ij(os.path.join(OUT, f"P_{lvl:02d}.IJ.out.00000"), L["n"], L["nc"], PI, PJ, PA)
    print("levels", [L["n"] for L in levels])


if __name__ == "__main__":
    main()
