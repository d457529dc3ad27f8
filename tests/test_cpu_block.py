"""The oracle's multi-RHS block Jacobi PCG (or_pcg_block: M systems on one pattern, dot products
in the library's block order), pinned without the library:
  - every member's solution solves its own system (scipy spsolve, relative error <= 1e-9);
  - every member takes the iteration count of the single-system PCG of the same system (or_pcg:
    the same recurrences, only the dot-product rounding order differs) and lands within 1e-10 of
    its solution;
  - a member with a zero right-hand side stops at once with x = 0; members padded to a power of
    two do not change the others; the member count need not be a power of two."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from oracle import pyoracle as O


def _laplace(nx, ny, shift, seed):
    rng = np.random.default_rng(seed)
    n = nx * ny
    idx = np.arange(n).reshape(ny, nx)
    rows, cols, vals = [], [], []
    for (a, b) in ((idx[:, :-1], idx[:, 1:]), (idx[:-1, :], idx[1:, :])):
        w = rng.uniform(0.5, 2.0, a.size)
        rows += [a.ravel(), b.ravel()]
        cols += [b.ravel(), a.ravel()]
        vals += [-w, -w]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A = A + sp.diags(np.abs(A).sum(1).A1 + shift)
    A = A.tocsr()
    A.sort_indices()
    return A


def _members(k, nx=40, ny=25):
    base = _laplace(nx, ny, 0.05, 1)
    vals, bs = [], []
    for m in range(k):
        A = _laplace(nx, ny, 0.05 + 0.01 * m, 10 + m)
        assert np.array_equal(A.indptr, base.indptr) and np.array_equal(A.indices, base.indices)
        vals.append(A.data.copy())
        bs.append(np.random.default_rng(100 + m).normal(size=A.shape[0]))
    return base, np.array(vals), np.array(bs)


def test_block_members_solve_their_systems():
    base, vals, bs = _members(6)
    x, it, rr = O.pcg_block(base.indptr, base.indices, vals, bs, rtol=1e-12, maxiter=3000, nthreads=4)
    for m in range(vals.shape[0]):
        A = sp.csr_matrix((vals[m], base.indices, base.indptr), shape=base.shape)
        xs = sla.spsolve(A.tocsc(), bs[m])
        assert np.linalg.norm(x[m] - xs) / np.linalg.norm(xs) < 1e-9, m
        xo, ito, _, _ = O.pcg(base.indptr, base.indices, vals[m], bs[m], rtol=1e-12, maxiter=3000)
        assert it[m] == ito, (m, it[m], ito)
        assert np.linalg.norm(x[m] - xo) / np.linalg.norm(xo) < 1e-10, m


def test_block_zero_rhs_padding_and_widths():
    base, vals, bs = _members(5)
    bs[2] = 0.0
    x5, it5, _ = O.pcg_block(base.indptr, base.indices, vals, bs, rtol=1e-12, maxiter=3000, nthreads=4)
    assert it5[2] == 0 and not x5[2].any()
    # members 0, 1 alone in a block of two: other solutions, same iteration counts and within 1e-10
    x2, it2, _ = O.pcg_block(base.indptr, base.indices, vals[:2], bs[:2], rtol=1e-12, maxiter=3000)
    for m in range(2):
        assert it2[m] == it5[m]
        assert np.linalg.norm(x2[m] - x5[m]) / np.linalg.norm(x5[m]) < 1e-10
    # the five-member block is padded to eight: the same bits as an explicit eight-member block
    # with unit-valued, zero right-hand-side members 5-7
    vals8 = np.vstack([vals, np.ones((3, vals.shape[1]))])
    bs8 = np.vstack([bs, np.zeros((3, bs.shape[1]))])
    x8, it8, _ = O.pcg_block(base.indptr, base.indices, vals8, bs8, rtol=1e-12, maxiter=3000)
    assert np.array_equal(x8[:5], x5) and np.array_equal(it8[:5], it5)
    assert (it8[5:] == 0).all()
