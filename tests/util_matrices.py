"""Random CSR inputs shared by the CPU and GPU tests."""
import numpy as np
import scipy.sparse as sp


def rand_csr(m, n, density, seed, shuffle_rows=True, empty_rows=()):
    """Random CSR with unsorted columns inside rows (HYPRE rows are not sorted)."""
    rng = np.random.default_rng(seed)
    S = sp.random(m, n, density=density, random_state=rng, format="csr")
    S.data = rng.standard_normal(S.nnz)
    S = S.tolil()
    for r in empty_rows:
        S.rows[r] = []
        S.data[r] = []
    S = S.tocsr()
    I, J, A = S.indptr.astype(np.int32), S.indices.astype(np.int32).copy(), S.data.copy()
    if shuffle_rows:
        for r in range(m):
            p = rng.permutation(I[r + 1] - I[r]) + I[r]
            J[I[r]:I[r + 1]] = J[p]
            A[I[r]:I[r + 1]] = A[p]
    return I, J, A


def to_scipy(m, n, I, J, A):
    return sp.csr_matrix((A, J, I), shape=(m, n))


def dense_rows_csr(m, n, per_row, seed):
    """Rows with exactly per_row entries in random column order."""
    rng = np.random.default_rng(seed)
    J = np.concatenate([rng.permutation(n)[:per_row] for _ in range(m)]).astype(np.int32)
    I = (np.arange(m + 1) * per_row).astype(np.int32)
    return I, J, rng.standard_normal(m * per_row)
