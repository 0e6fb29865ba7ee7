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


def skewed_csr(m, n, long_block_rows, long_len, short_len, seed, empty_rows=()):
    """Mostly short rows plus a few block rows (4 CSR rows each) of long_len
    entries per row: blocks per block row spread far above 12 (the reference's
    switch to its balanced SpMV)."""
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(m):
        if r in empty_rows:
            k = 0
        elif r // 4 in long_block_rows:
            k = long_len
        else:
            k = short_len
        rows.append(np.sort(rng.permutation(n)[:k]))
    I = np.zeros(m + 1, np.int32)
    I[1:] = np.cumsum([len(c) for c in rows])
    J = np.concatenate(rows).astype(np.int32)
    return I, J, rng.standard_normal(len(J))


def block_count_csr(counts, n, seed):
    """Block row br holds exactly counts[br] blocks (one entry each, on its
    first CSR row), so block totals hit chosen tile edges."""
    rng = np.random.default_rng(seed)
    bcol = n // 4
    m = 4 * len(counts)
    I = np.zeros(m + 1, np.int32)
    cols = []
    for br, c in enumerate(counts):
        k = np.sort(rng.permutation(bcol)[:c])
        cols.append(4 * k + rng.integers(0, 4, c))
        I[4 * br + 1:4 * br + 5] = I[4 * br] + c
    J = np.concatenate(cols).astype(np.int32)
    return m, I, J, rng.standard_normal(len(J))
