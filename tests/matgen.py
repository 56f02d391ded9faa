"""Small seeded sparse matrices for the parity tests (numpy, CPU only).

Each generator returns (m, n, rowptr, colidx, val) with int32 CSR arrays and
float64 values, rows in ascending column order unless `shuffle_rows` is set
(the reference keeps file order inside a row, so unsorted rows are legal).
"""
import numpy as np


def _csr_from_lists(m, n, rows, rng, vals="uniform"):
    lens = np.array([len(r) for r in rows], dtype=np.int64)
    rowptr = np.zeros(m + 1, np.int64)
    np.cumsum(lens, out=rowptr[1:])
    colidx = (np.concatenate(rows) if rowptr[-1] else np.zeros(0)).astype(np.int32)
    nnz = int(rowptr[-1])
    if vals == "uniform":
        val = rng.uniform(-1.0, 1.0, nnz)
    elif vals == "normal":
        val = rng.standard_normal(nnz)
    else:
        val = np.ones(nnz)
    return m, n, rowptr.astype(np.int32), colidx, val.astype(np.float64)


def random_csr(m, n, density, seed, shuffle_rows=False, vals="uniform"):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(m):
        k = rng.binomial(n, density)
        c = np.sort(rng.choice(n, size=k, replace=False))
        if shuffle_rows:
            rng.shuffle(c)
        rows.append(c)
    return _csr_from_lists(m, n, rows, rng, vals)


def powerlaw_csr(m, n, seed, zipf=2.0, max_len=64, run_frac=0.3, run=8):
    """Power-law row lengths, uniform columns, some rows as contiguous runs."""
    rng = np.random.default_rng(seed)
    ks = np.arange(1, max_len + 1, dtype=np.float64)
    p = ks ** -zipf
    p /= p.sum()
    lens = rng.choice(ks.astype(np.int64), size=m, p=p)
    rows = []
    for i in range(m):
        k = min(int(lens[i]), n)
        if rng.random() < run_frac and n >= run:
            s = int(rng.integers(0, n - run + 1))
            c = np.arange(s, s + run)
            extra = rng.choice(n, size=max(0, k - run), replace=False)
            c = np.unique(np.concatenate([c, extra]))
        else:
            c = np.sort(rng.choice(n, size=k, replace=False))
        rows.append(c)
    return _csr_from_lists(m, n, rows, rng)


def dense_column_csr(m, n, seed, dense_cols=3, dense_len=200):
    """Mostly sparse, plus a few very long columns (forces balance splits)."""
    rng = np.random.default_rng(seed)
    sets = [set(rng.choice(n, size=int(rng.integers(0, 4)), replace=False).tolist())
            for _ in range(m)]
    for j in rng.choice(n, size=dense_cols, replace=False):
        for i in rng.choice(m, size=min(dense_len, m), replace=False):
            sets[i].add(int(j))
    rows = [np.array(sorted(s), dtype=np.int64) for s in sets]
    return _csr_from_lists(m, n, rows, rng)


def equal_rows_csr(m, n, seed):
    """Rows whose columns repeat across 32 consecutive sorted columns at the
    same depth, so tiles of format 0 and 2 appear."""
    rng = np.random.default_rng(seed)
    rows = [np.zeros(0, np.int64) for _ in range(m)]
    heavy = rng.choice(m, size=max(1, m // 16), replace=False)
    for i in heavy:
        rows[i] = np.arange(n, dtype=np.int64)
    for i in range(m):
        if i not in set(heavy.tolist()):
            rows[i] = np.sort(rng.choice(n, size=int(rng.integers(0, 3)), replace=False))
    return _csr_from_lists(m, n, rows, rng)


def banded_mixed_csr(m, n, seed, bw=64, lo=8, hi=32, zipf=1.8, max_len=200, offset=0,
                     mixed=True):
    """Even rows banded around column i + offset (bandwidth bw, lo..hi
    entries), odd rows Zipf lengths with uniform columns (mixed) or banded too:
    the small form of the 10^6-row benchmark matrix."""
    rng = np.random.default_rng(seed)
    ks = np.arange(1, max_len + 1, dtype=np.float64)
    p = ks ** -zipf
    p /= p.sum()
    rows = []
    for i in range(m):
        if i % 2 == 0 or not mixed:
            c0, c1 = max(0, i + offset - bw), min(n, i + offset + bw + 1)
            k = min(int(rng.integers(lo, hi + 1)), max(0, c1 - c0))
            c = np.sort(rng.choice(np.arange(c0, c1), size=k, replace=False)) if k else \
                np.zeros(0, np.int64)
        else:
            k = min(int(rng.choice(ks.astype(np.int64), p=p)), n)
            c = np.sort(rng.choice(n, size=k, replace=False))
        rows.append(c)
    return _csr_from_lists(m, n, rows, rng, vals="normal")


def csr_to_dense(m, n, rowptr, colidx, val):
    A = np.zeros((m, n))
    for i in range(m):
        for e in range(rowptr[i], rowptr[i + 1]):
            A[i, colidx[e]] += val[e]
    return A
