"""The HYPRE-side hooks compiled as C (tests/hypre_hooks/hypre_hooks.c: the
bodies INTEGRATION.md shows for hypreDevice_CSRSpGemm and
hypre_CSRMatrixMatvecDevice2, over stand-ins of the HYPRE types) and run in a
fresh process on the GPU; their outputs against the oracle:

* C = A*A through the SpGEMM hook: the row pattern equals HYPRE's host
  SpGEMM's (as column sets), values within 1e-12 of |A||A|; with A
  pattern-only the hook returns the same pattern and no values;
* y = 2.5 A x - 0.5 y (b aliasing y, as the hook passes it) and z = A^T x
  through the SpMV hook, within 1e-12 of the |A||x| row scale."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "hypre_hooks", "hypre_hooks")


def operator(nx, ny, nz):
    """The driver's matrix: 7-point, diagonal 6 + 0.01 (r % 5) first,
    off-diagonals -(1 + 0.001 ((r + 2 c) % 11)) in ascending order."""
    n = nx * ny * nz
    rows, cols, vals = [], [], []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                r = (z * ny + y) * nx + x
                rows.append(r), cols.append(r), vals.append(6.0 + 0.01 * (r % 5))
                for ok, c in ((z > 0, r - nx * ny), (y > 0, r - nx), (x > 0, r - 1),
                              (x < nx - 1, r + 1), (y < ny - 1, r + nx), (z < nz - 1, r + nx * ny)):
                    if ok:
                        rows.append(r), cols.append(c), vals.append(-(1.0 + 0.001 * ((r + 2 * c) % 11)))
    I = np.zeros(n + 1, np.int32)
    np.add.at(I, np.array(rows) + 1, 1)
    return np.cumsum(I).astype(np.int32), np.array(cols, np.int32), np.array(vals)


@pytest.mark.gpu
def test_compiled_hooks_match_oracle(tmp_path):
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    nx, ny, nz = 12, 10, 9
    print(f"[hooks] {EXE} {nx} {ny} {nz}", flush=True)
    r = subprocess.run([EXE, str(nx), str(ny), str(nz), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    n = nx * ny * nz
    I, J, V = operator(nx, ny, nz)
    rd = lambda name, dt: np.fromfile(os.path.join(tmp_path, name), dtype=dt)  # noqa: E731
    CI, CJ, CA = rd("C_i.i32", np.int32), rd("C_j.i32", np.int32), rd("C_data.f64", np.float64)
    RI, RJ, RA = O.csr_spgemm(n, n, I, J, V, I, J, V)
    assert np.array_equal(CI, RI)
    A = sp.csr_matrix((V, J, I), shape=(n, n))
    scale = (abs(A) @ abs(A)).tocsr()
    for row in range(n):
        a, b = slice(CI[row], CI[row + 1]), slice(RI[row], RI[row + 1])
        oc, orf = np.argsort(CJ[a]), np.argsort(RJ[b])
        assert np.array_equal(CJ[a][oc], RJ[b][orf]), row
        sc = np.asarray(scale[row, RJ[b][orf]].todense()).ravel()
        assert np.all(np.abs(CA[a][oc] - RA[b][orf]) <= 1e-12 * sc), row
    assert np.array_equal(rd("Cp_j.i32", np.int32), CJ)
    x = 1.0 / (1.0 + np.arange(n) % 13)
    y0 = 0.25 * (np.arange(n) % 7) - 0.5
    y, z = rd("y.f64", np.float64), rd("z.f64", np.float64)
    ey = O.csr_matvec(n, I, J, V, 2.5, x, -0.5, y0)
    ez = O.csr_matvec_t(n, n, I, J, V, 1.0, x)
    sy = 2.5 * (abs(A) @ np.abs(x)) + 0.5 * np.abs(y0)
    sz = abs(A).T @ np.abs(x)
    assert np.all(np.abs(y - ey) <= 1e-12 * sy)
    assert np.all(np.abs(z - ez) <= 1e-12 * sz)


def test_hooks_binary_is_built_and_linked():
    """CPU: the hooks compile against include/amgt_mbsr.h and link the in-tree
    library (no device call here)."""
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    lib = [ln for ln in out.splitlines() if "libamgt_mbsr.so" in ln]
    assert lib, out
    path = lib[0].split("=>")[1].split("(")[0].strip()
    assert os.path.realpath(path) == os.path.realpath(
        os.path.join(ROOT, "amgt_amd", "libamgt_mbsr.so")), out


def test_integration_doc_quotes_the_compiled_hooks():
    """CPU: every hook block of hypre_hooks.c appears verbatim in INTEGRATION.md,
    so the documented hooks are the compiled ones."""
    src = open(os.path.join(ROOT, "tests", "hypre_hooks", "hypre_hooks.c")).read()
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = src.split("/* ---- hook: ")[1:]
    assert len(blocks) == 3
    for blk in blocks:
        body = blk.split("\n", 1)[1].split("/* ---- end hook ---- */")[0].rstrip("\n")
        assert body in doc, body[:80]
