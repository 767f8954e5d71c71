"""Hypergraph-product codes (CPU): construction and the logical operators
that replace the dense row basis in the success test.  Pinned by GF(2)
linear algebra on the dense matrices (numpy), not by the construction's own
formulas, and against the oracle's row-space membership."""
import numpy as np
import pytest

import joint_bp_plus_pp_amd as jbp


def dense(code, side):
    off, ev = code.csr(side)
    H = np.zeros((off.size - 1, code.n), np.uint8)
    for c in range(off.size - 1):
        H[c, ev[off[c]:off[c + 1]]] ^= 1
    return H


def rank2(A):
    A = (A % 2).astype(np.uint8)
    r = 0
    rows, cols = A.shape
    for c in range(cols):
        if r == rows:
            break
        piv = np.nonzero(A[r:, c])[0]
        if piv.size == 0:
            continue
        p = r + piv[0]
        A[[r, p]] = A[[p, r]]
        nz = np.nonzero(A[:, c])[0]
        A[nz[nz != r]] ^= A[r]
        r += 1
    return r


def logical_matrix(code, side):
    off, q = code.logicals(side)
    L = np.zeros((off.size - 1, code.n), np.uint8)
    for i in range(off.size - 1):
        L[i, q[off[i]:off[i + 1]]] = 1
    return L


@pytest.mark.parametrize("m,n,seed", [(6, 8, 5), (9, 12, 5), (12, 16, 3)])
def test_logicals_span_the_quotient(m, n, seed):
    # k = n_q - rank Hx - rank Hz rows per side; each row in ker of the other
    # side's matrix; rows independent modulo the same side's row space.
    code = jbp.Code.hgp(m, n, 3, 4, seed=seed)
    assert code.n == n * n + m * m
    Hx, Hz = dense(code, 0), dense(code, 1)
    assert not ((Hx.astype(int) @ Hz.T.astype(int)) % 2).any()
    k = code.n - rank2(Hx) - rank2(Hz)
    for side, other, same in ((0, Hz, Hx), (1, Hx, Hz)):
        L = logical_matrix(code, side)
        assert L.shape[0] == k
        assert not ((other.astype(int) @ L.T.astype(int)) % 2).any()
        assert rank2(np.vstack([same, L])) - rank2(same) == k


def test_logical_test_is_row_space_membership(oracle):
    # r = a.Hz + b.L1 has zero X syndrome; it is in rowspace(Hz) iff b = 0,
    # and the side-0 parities must say the same as the oracle's elimination.
    code = jbp.Code.hgp(9, 12, 3, 4, seed=5)
    Hz = dense(code, 1)
    L0, L1 = logical_matrix(code, 0), logical_matrix(code, 1)
    x_checks = code.checks(0)
    z_checks = code.checks(1)
    oc = oracle.code_from_checks(code.n, x_checks, z_checks)
    rng = np.random.default_rng(7)
    seen = set()
    for trial in range(60):
        a = rng.integers(0, 2, Hz.shape[0])
        b = rng.integers(0, 2, L1.shape[0]) if trial % 2 else np.zeros(L1.shape[0], int)
        r = (a @ Hz + b @ L1) % 2
        by_logicals = not ((L0.astype(int) @ r) % 2).any()
        assert by_logicals == oc.in_row_space(1, r) == (not b.any())
        seen.add(by_logicals)
    assert seen == {True, False}


def test_full_size_hgp_code():
    # HGP of a seeded (3,4)-regular 150 x 200 matrix: n = 62,500, 30,000
    # checks and 210,000 edges per side, check degree 7, joint variable
    # degree 6 (n x n block) or 8 (m x m block).
    code = jbp.Code.hgp(150, 200, 3, 4, seed=1)
    assert (code.n, code.mx, code.mz) == (62500, 30000, 30000)
    assert code.x_edges == code.z_edges == 210000
    for side in (0, 1):
        off, ev = code.csr(side)
        assert np.all(np.diff(off) == 7)
        deg = np.bincount(ev, minlength=code.n)
        assert np.all(deg[:40000] == 3) and np.all(deg[40000:] == 4)
    rows = [logical_matrix_rows(code, s) for s in (0, 1)]
    assert rows[0] == rows[1]
    # (n - r)^2 + (m - r)^2 logicals for rank r of the classical matrix
    assert any((200 - r) ** 2 + (150 - r) ** 2 == rows[0] for r in range(151))


def logical_matrix_rows(code, side):
    off, _ = code.logicals(side)
    return off.size - 1


def test_params_code_has_no_product_logicals():
    from conftest import PARAMS
    code = jbp.Code.from_params(PARAMS)
    assert code.logicals(0) is None and code.logicals(1) is None
