"""The tournament-pivoting restatement (oracle/lu.py) pinned on the CPU: one leaf is GEPP (LAPACK's pivots),
dominant rows are always elected whatever the tree's shape, and the factor satisfies P A = L U."""
import numpy as np
import pytest
import scipy.linalg as sla

from oracle import lu as olu
from tests.cases import seeded_matrix


def test_single_leaf_is_lapack_partial_pivoting():
    """With every row of the panel in one leaf the tournament is GEPP of the panel: its interchanges are the first
    64 of LAPACK's on the whole matrix (partial pivoting's choices on the first panel do not depend on blocking)."""
    A = seeded_matrix(200, swap=True).T.copy()
    win = olu.tslu_winners(A, 0, 64, RL=256, AR=4)
    ipiv = olu.interchanges(win, 0, np.arange(200))
    _, piv = sla.lu_factor(A)
    assert np.array_equal(ipiv, piv[:64])


@pytest.mark.parametrize('RL,AR,n', [(128, 2, 1000), (256, 4, 1300), (64, 3, 700), (128, 2, 130)])
def test_dominant_rows_win_in_column_order(RL, AR, n):
    """One row per panel column is 1e6 times larger in that column than anything else: every GEPP of any node
    picks it first among the rows it sees, so the winners are exactly those rows in column order - a dropped
    child, a wrong stacking or a lost nominee would fail this."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    jb = 64
    dom = rng.choice(n, jb, replace=False)
    for k, r in enumerate(dom):
        A[r, k] = 1e6 * (1 + 1j)
    win = olu.tslu_winners(A, 0, jb, RL, AR)
    assert np.array_equal(win, dom)


@pytest.mark.parametrize('n,RL,AR', [(300, 128, 2), (517, 256, 4), (70, 128, 2)])
def test_factor_is_p_a_equals_l_u(n, RL, AR):
    A = seeded_matrix(n).T.copy()
    LU, ipiv = olu.getrf_tournament(A, RL=RL, AR=AR)
    assert np.all(ipiv >= np.arange(n)) and np.all(ipiv < n)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    perm = olu.perm_of(ipiv)
    assert np.abs(A[perm] - L @ U).max() <= 1e-12 * np.abs(A).max()
    assert np.abs(L).max() < 10                       # pivoting kept the multipliers small
