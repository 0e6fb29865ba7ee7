"""Host matrix readers (amgt_read_ij / amgt_read_mm), CPU only.

Pins: scipy.io.mmread on Matrix Market files (general, symmetric, integer);
hand-computed IJ semantics of hypre_IJMatrixRead (IJ_mv/IJMatrix.c:112-242:
own rows set, the last value wins; other rows added to the owner's value);
round trips through the IJ writer over several files; and, where the reference
tree is present, HYPRE's own 4-file IJ fixture (test/TEST_ams/mfem.A.*)
against a plain parse of its lines."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from amgt_amd import _lib as L
from amgt_amd import matrix_io as MIO
from tests.util_matrices import rand_csr


def sorted_csr(S):
    S = sp.csr_matrix(S)
    S.sort_indices()
    return S


def same(nr, nc, I, J, V, S):
    S = sorted_csr(S)
    assert (nr, nc) == S.shape
    assert np.array_equal(I, S.indptr) and np.array_equal(J, S.indices)
    assert np.array_equal(V, S.data)


def test_mm_general_symmetric_integer(tmp_path):
    I, J, A = rand_csr(37, 53, 0.15, 1)
    S = sp.csr_matrix((A, J, I), shape=(37, 53))
    p = str(tmp_path / "g.mtx")
    scipy.io.mmwrite(p, S, precision=17)
    same(*MIO.read_mm(p), scipy.io.mmread(p))
    T = sp.random(40, 40, density=0.1, random_state=2, format="csr")
    T = T + T.T + sp.eye(40)
    p2 = str(tmp_path / "s.mtx")
    scipy.io.mmwrite(p2, T, symmetry="symmetric", precision=17)
    same(*MIO.read_mm(p2), scipy.io.mmread(p2))
    R = sp.csr_matrix((np.arange(1, 8, dtype=np.int64), ([0, 1, 2, 3, 4, 4, 0], [0, 1, 2, 3, 4, 0, 4])),
                      shape=(5, 5))
    p3 = str(tmp_path / "i.mtx")
    scipy.io.mmwrite(p3, R, field="integer")
    same(*MIO.read_mm(p3), scipy.io.mmread(p3).astype(np.float64))


def test_mm_errors(tmp_path):
    p = tmp_path / "pat.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n")
    with pytest.raises(L.AmgtError) as e:
        MIO.read_mm(str(p))
    assert e.value.status == L.AMGT_ERR_UNSUPPORTED
    p.write_text("not a matrix\n")
    with pytest.raises(L.AmgtError) as e:
        MIO.read_mm(str(p))
    assert e.value.status == L.AMGT_ERR_INPUT
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(L.AmgtError) as e:
        MIO.read_mm(str(p))
    assert e.value.status == L.AMGT_ERR_INPUT
    with pytest.raises(L.AmgtError):
        MIO.read_mm(str(tmp_path / "missing.mtx"))


def test_ij_semantics(tmp_path):
    pre = str(tmp_path / "A")
    with open(pre + ".00000", "w") as f:
        f.write("0 1 0 3\n")
        f.write("0 0 4.0\n0 1 -1.0\n1 1 5.0\n")
        f.write("0 0 6.0\n")        # own row, set again: replaces 4.0
        f.write("3 3 0.25\n")       # row 3 belongs to file 1: added
        f.write("2 0 1.5\n")        # row 2 belongs to file 1: added, no set there
    with open(pre + ".00001", "w") as f:
        f.write("2 3 0 3\n")
        f.write("3 3 2.0\n2 2 7.0\n3 2 -1.0\n")
    nr, nc, I, J, V = MIO.read_ij(pre, 2)
    D = np.zeros((4, 4))
    D[0, 0], D[0, 1], D[1, 1] = 6.0, -1.0, 5.0
    D[3, 3] = 2.0 + 0.25
    D[2, 0] = 1.5
    D[2, 2], D[3, 2] = 7.0, -1.0
    same(nr, nc, I, J, V, sp.csr_matrix(D))


@pytest.mark.parametrize("nfiles", [1, 3])
def test_ij_round_trip(tmp_path, nfiles):
    I, J, A = rand_csr(61, 47, 0.12, 7)
    pre = str(tmp_path / "M")
    MIO.write_ij(pre, 61, 47, I, J, A, nfiles)
    same(*MIO.read_ij(pre, nfiles), sp.csr_matrix((A, J, I), shape=(61, 47)))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_AMS = os.path.join(GOLDEN, "ij_ams", "mfem.Aalpha")


def test_reference_ij_fixture():
    """HYPRE's own IJ matrix (TEST_ams/mfem.Aalpha, 4 rank files, copied from
    the reference's test data) against a plain parse of the lines."""
    rows, cols, vals, hi_r, hi_c = [], [], [], 0, 0
    for f in range(4):
        with open(f"{REF_AMS}.{f:05d}") as fh:
            ilo, ihi, jlo, jhi = map(int, fh.readline().split())
            hi_r, hi_c = max(hi_r, ihi), max(hi_c, jhi)
            d = np.loadtxt(fh, ndmin=2)
        rows.append(d[:, 0].astype(np.int64))
        cols.append(d[:, 1].astype(np.int64))
        vals.append(d[:, 2])
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    assert len(np.unique(r * (hi_c + 1) + c)) == len(r)  # no repeated entries in the fixture
    S = sp.csr_matrix((v, (r, c)), shape=(hi_r + 1, hi_c + 1))
    same(*MIO.read_ij(REF_AMS, 4), S)


def load_level_fixture(tmp_path, name):
    """A DEBUG_SAVE_ALL_OPS file of the fsai job's hierarchy: (n_rows, n_cols,
    rows, cols, values) in the stored order, and the library reader's CSR."""
    import gzip
    import shutil
    src = os.path.join(GOLDEN, "hierarchy_ij_fsai103", name + ".IJ.out.00000.gz")
    dst = tmp_path / (name + ".IJ.out.00000")
    with gzip.open(src, "rb") as fi, open(dst, "wb") as fo:
        shutil.copyfileobj(fi, fo)
    with open(dst) as fh:
        ilo, ihi, jlo, jhi = map(int, fh.readline().split())
        d = np.loadtxt(fh, ndmin=2)
    raw = (ihi + 1, jhi + 1, d[:, 0].astype(np.int64), d[:, 1].astype(np.int64), d[:, 2])
    return raw, MIO.read_ij(str(tmp_path / (name + ".IJ.out")), 1)


def test_hierarchy_fixtures_match_oracle(tmp_path):
    """The committed per-level operators (HYPRE's DEBUG_SAVE_ALL_OPS layout,
    tests/golden/make_hierarchy_fixtures.py) against the oracle's host-order
    hierarchy of the same job: entries in HYPRE's stored order, values to the
    14 printed decimals; read back by amgt_read_ij."""
    from oracle import oracle as O
    I, J, A = O.stencil(7, 10, 10, 10)
    levels = O.amg_hierarchy(1000, I, J, A, theta=0.25, max_row_sum=1.0, max_elmts=4,
                             max_coarse_size=9, coarsen="hmis", rap="kt")
    assert [L_["n"] for L_ in levels] == [1000, 500, 99, 14, 4]
    for lvl, L_ in enumerate(levels):
        mats = [("A", L_["n"], L_["I"], L_["J"], L_["A"])]
        if "P" in L_:
            mats.append(("P", L_["nc"], *L_["P"]))
        for tag, ncols, MI, MJ, MV in mats:
            (nr, nc, r, c, v), rd = load_level_fixture(tmp_path, f"{tag}_{lvl:02d}")
            assert (nr, nc) == (L_["n"], ncols)
            assert np.array_equal(r, np.repeat(np.arange(L_["n"]), np.diff(MI)))
            assert np.array_equal(c, MJ)
            assert np.all(np.abs(v - MV) <= 1e-14 * np.abs(MV))
            same(*rd, sp.csr_matrix((v, (r, c)), shape=(nr, nc)))
