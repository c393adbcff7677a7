"""Tournament pivoting (TSLU, the panel factorisation of CALU: Grigori, Demmel and Xiang, SIAM J. Matrix Anal.
Appl. 32 (2011) 1317-1350) restated in NumPy to check the pivots of the GPU's tournament-pivoting LU.
TEST INFRASTRUCTURE ONLY.

The reference factors with LAPACK's partial pivoting (src/mom25d/fast.py:318-320: scipy lu_factor(Z.T);
src/mom25d/gpu_vk.py:515-633 on its GPU backend); the GPU's partial-pivoting path is checked against LAPACK
directly. The tournament path is this build's own: a panel of TW columns takes its pivot rows from a tree of
GEPPs - leaves of RL rows, nodes of up to AR children whose nominees (original rows, stacked in child order) are
factored again - and the rest of the factorisation is the usual blocked right-looking update. Pinned by: a single
leaf reduces to GEPP (LAPACK's pivots, tests/test_oracle_lu.py), a panel with one dominant row per column must
elect exactly those rows in column order, and P A = L U.
"""
import numpy as np


def gepp_order(M):
    """The pivot order of GEPP on M (r x c, complex): indices into M's rows, min(r, c) of them. LAPACK's rule: the
    first largest |Re| + |Im| in the current order; a pivot row swaps places with the row at the step's position."""
    M = np.array(M, complex)
    r, c = M.shape
    rows = list(range(r))
    for k in range(min(r, c)):
        col = np.abs(M[k:, k].real) + np.abs(M[k:, k].imag)
        p = k + int(np.argmax(col))
        if p != k:
            M[[k, p]] = M[[p, k]]
            rows[k], rows[p] = rows[p], rows[k]
        if M[k, k] != 0:                                   # times the reciprocal conj(p) / |p|^2, as the GPU's GEPP
            p = M[k, k]
            M[k + 1:, k] *= np.conj(p) * (1.0 / (p.real * p.real + p.imag * p.imag))
        M[k + 1:, k + 1:] -= np.outer(M[k + 1:, k], M[k, k + 1:])
    return rows[:min(r, c)]


def tslu_winners(A, j0, jb, RL, AR):
    """The panel's pivot rows (global row indices of A, A[row, col], in pivot order) elected by the tournament over
    rows [j0, n) of columns [j0, j0 + jb)."""
    n = A.shape[0]
    cand = []
    for r0 in range(j0, n, RL):
        rows = np.arange(r0, min(n, r0 + RL))
        cand.append(rows[gepp_order(A[rows, j0:j0 + jb])])
    while len(cand) > 1:
        nxt = []
        for q in range(0, len(cand), AR):
            group = cand[q:q + AR]
            if len(group) == 1:
                nxt.append(group[0])
                continue
            rows = np.concatenate(group)
            nxt.append(rows[gepp_order(A[rows, j0:j0 + jb])])
        cand = nxt
    return cand[0]


def interchanges(win, j0, perm):
    """LAPACK-style interchanges bringing win[k] to position j0 + k in order; perm (position -> row) is updated in
    place. Returns ipiv[j0 : j0 + len(win)]."""
    out = []
    for k, w in enumerate(win):
        t = j0 + k
        p = int(np.nonzero(perm == w)[0][0])
        out.append(p)
        perm[t], perm[p] = perm[p], perm[t]
    return np.array(out)


def getrf_tournament(A, TW=64, RL=128, AR=2):
    """(LU, ipiv) of A (A[row, col], complex128) by tournament pivoting, blocked right-looking, in FP64: the
    layout of LAPACK's getrf (L unit lower and U in place, ipiv 0-based interchanges)."""
    A = np.array(A, complex)
    n = A.shape[0]
    ipiv = np.zeros(n, int)
    for j0 in range(0, n, TW):
        jb = min(TW, n - j0)
        win = tslu_winners(A, j0, jb, RL, AR)
        perm = np.arange(n)
        ipiv[j0:j0 + jb] = interchanges(win, j0, perm)
        A[j0:] = A[perm[j0:]]
        # the panel's top block has its pivots on the diagonal now: unpivoted LU of it
        T = A[j0:j0 + jb, j0:j0 + jb]
        for k in range(jb):
            p = T[k, k]
            if p != 0:
                T[k + 1:, k] *= np.conj(p) * (1.0 / (p.real * p.real + p.imag * p.imag))
            T[k + 1:, k + 1:] -= np.outer(T[k + 1:, k], T[k, k + 1:])
        U11 = np.triu(T)
        L11 = np.tril(T, -1) + np.eye(jb)
        A[j0 + jb:, j0:j0 + jb] = np.linalg.solve(U11.T, A[j0 + jb:, j0:j0 + jb].T).T
        A[j0:j0 + jb, j0 + jb:] = np.linalg.solve(L11, A[j0:j0 + jb, j0 + jb:])
        A[j0 + jb:, j0 + jb:] -= A[j0 + jb:, j0:j0 + jb] @ A[j0:j0 + jb, j0 + jb:]
    return A, ipiv


def perm_of(ipiv):
    """The row permutation (position -> original row) of LAPACK-style interchanges."""
    perm = np.arange(len(ipiv))
    for i, p in enumerate(ipiv):
        perm[i], perm[p] = perm[p], perm[i]
    return perm
