"""Block-diagonal sparse matrices (jax_cosmo/sparse.py:37-389), float64 host.

A sparse matrix is a [ny, nx, ndiag] array of ny x nx blocks, each block a
diagonal of length ndiag.
"""
import numpy as np

__all__ = ["is_sparse", "check_sparse", "to_dense", "dot", "sparse_dot_vec", "inv", "slogdet", "det"]


def is_sparse(sparse):  # sparse.py:37-39
    return np.asarray(sparse).ndim == 3


def check_sparse(sparse, square=False):  # sparse.py:42-49
    sparse = np.asarray(sparse)
    if sparse.ndim != 3:
        raise ValueError("Expected 3D array of sparse diagonals.")
    if square and (sparse.shape[0] != sparse.shape[1]):
        raise ValueError("Expected a square matrix.")
    return sparse


def to_dense(sparse):  # sparse.py:52-68
    sparse = check_sparse(sparse)
    a, b, nd = sparse.shape
    out = np.zeros((a * nd, b * nd), dtype=sparse.dtype)
    idx = np.arange(nd)
    for i in range(a):
        for j in range(b):
            out[i * nd + idx, j * nd + idx] = sparse[i, j]
    return out


def sparse_dot_vec(sparse, vec):  # sparse.py:140-162
    sparse = np.asarray(sparse)
    a, b, nd = sparse.shape
    return np.einsum("ijk,jk->ik", sparse, np.asarray(vec).reshape(b, nd)).reshape(-1)


def vec_dot_sparse(vec, sparse):  # sparse.py:187-209
    sparse = np.asarray(sparse)
    a, b, nd = sparse.shape
    return np.einsum("ik,ijk->jk", np.asarray(vec).reshape(a, nd), sparse).reshape(-1)


def sparse_dot_sparse(s1, s2):  # sparse.py:234-261
    return np.einsum("ijk,jlk->ilk", np.asarray(s1), np.asarray(s2))


def dot(*args):  # sparse.py:71-137 (vector/dense/sparse combinations)
    if len(args) == 2:
        A, B = (np.asarray(x) for x in args)
        if is_sparse(A) and is_sparse(B):
            return sparse_dot_sparse(A, B)
        if is_sparse(A):
            if B.ndim == 1:
                return sparse_dot_vec(A, B)
            return np.stack([sparse_dot_vec(A, B[:, c]) for c in range(B.shape[1])], axis=1)
        if is_sparse(B):
            if A.ndim == 1:
                return vec_dot_sparse(A, B)
            return np.stack([vec_dot_sparse(A[r], B) for r in range(A.shape[0])], axis=0)
        raise ValueError("at least one argument must be sparse")
    if len(args) == 3:
        X, Y, Z = (np.asarray(x) for x in args)
        if X.ndim != 2 or Y.ndim != 3 or Z.ndim != 2:
            raise ValueError("Can only handle dense @ sparse @ dense bilinear form.")
        return X @ dot(Y, Z)
    raise ValueError("Expected 2 or 3 input arrays but got %d." % len(args))


def inv(sparse):  # sparse.py:295-315: n_diag independent inverses
    sparse = check_sparse(sparse, square=True)
    return np.transpose(np.linalg.inv(np.transpose(sparse, (2, 0, 1))), (1, 2, 0))


def slogdet(sparse):  # sparse.py:322-366 (same value via per-diagonal blocks)
    sparse = check_sparse(sparse, square=True)
    s, ld = np.linalg.slogdet(np.transpose(sparse, (2, 0, 1)))
    return np.prod(s), np.sum(ld)


def det(sparse):  # sparse.py:369-389
    s, ld = slogdet(sparse)
    return s * np.exp(ld)
