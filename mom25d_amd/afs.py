"""The adaptive frequency sweep: a rational model of the S-parameters through as few EM solves as a tolerance allows.

The reference takes it from emsdev.afs (src/mom25d/afs.py:5-12; called by src/mom25d/run.py:69-73), a package not
available here; this is the build's own, from the published method:

  aaa             the AAA algorithm (Y. Nakatsukasa, O. Sete, L. N. Trefethen, "The AAA algorithm for rational
                  approximation", SIAM J. Sci. Comput. 40 (2018) A1494-A1522) for many functions sharing one
                  barycentric denominator (the set-valued form: every S entry interpolated with the same support
                  points and weights, chosen greedily where the worst entry is worst)
  adaptive_sweep  samples added one solve at a time where the last two models disagree most; done when two
                  successive models agree within tol everywhere on the grid and the model before the last sample
                  predicted that sample within tol (a model that stopped changing but does not follow new samples
                  is not converged)
"""
import numpy as np
import scipy.linalg as sla


class Barycentric:
    """r(z) = sum_k w_k F_k / (z - z_k) / sum_k w_k / (z - z_k) for support points z_k, values F_k (n, nf)."""

    def __init__(self, zk, Fk, wk):
        self.zk, self.Fk, self.wk = zk, Fk, wk

    def __call__(self, z):
        z = np.atleast_1d(np.asarray(z, complex))
        if not len(self.zk):
            return np.zeros((len(z), self.Fk.shape[1]), complex)
        d = z[:, None] - self.zk[None, :]
        hit = d == 0
        d[hit] = 1.0
        C = 1.0 / d
        out = (C @ (self.wk[:, None] * self.Fk)) / (C @ self.wk)[:, None]
        rows, cols = np.nonzero(hit)                   # at a support point: its value
        out[rows] = self.Fk[cols]
        return out

    def poles(self):
        """The approximant's poles: the finite eigenvalues of the barycentric arrowhead pencil."""
        n = len(self.zk)
        if n < 2:
            return np.zeros(0, complex)
        B = np.eye(n + 1, dtype=complex)
        B[0, 0] = 0
        E = np.zeros((n + 1, n + 1), complex)
        E[0, 1:], E[1:, 0] = self.wk, 1.0
        E[1:, 1:] = np.diag(self.zk)
        ev = sla.eigvals(E, B)
        return ev[np.isfinite(ev)]


def aaa(z, F, tol=1e-13, mmax=100):
    """The AAA approximant of samples F (m samples x nf functions, or m) at points z (m): support points added where
    the largest residual over the functions (each scaled by its largest sample) is, weights the last right singular
    vector of the stacked Loewner matrices, until that residual is within tol or mmax points are used."""
    z = np.asarray(z, complex).ravel()
    F = np.asarray(F, complex).reshape(len(z), -1)
    m, nf = F.shape
    scale = np.abs(F).max(axis=0)
    scale[scale == 0] = 1.0
    Fs = F / scale
    free = np.ones(m, bool)
    J = []
    R = np.broadcast_to(Fs.mean(axis=0), Fs.shape).copy()
    w = np.zeros(0, complex)
    for _ in range(min(mmax, m)):
        res = np.abs(Fs - R).max(axis=1)
        res[~free] = 0.0
        j = int(np.argmax(res))
        if res[j] <= tol:
            break
        J.append(j)
        free[j] = False
        zj, Fj = z[J], Fs[J]
        C = 1.0 / (z[free, None] - zj[None, :])
        if not free.any():
            w = np.ones(len(J), complex)
            break
        L = np.concatenate([(Fs[free, q][:, None] - Fj[None, :, q]) * C for q in range(nf)], axis=0)
        _, _, Vh = np.linalg.svd(L, full_matrices=len(J) > L.shape[0])
        w = Vh[-1].conj()
        R = Fs.copy()
        R[free] = (C @ (w[:, None] * Fj)) / (C @ w)[:, None]
    return Barycentric(z[J], F[J] if J else F[:0], w)


def _is_geometric(f):
    f = np.asarray(f, float)
    return len(f) > 2 and f[0] > 0 and np.allclose(np.diff(np.log(f)), np.log(f[1] / f[0]), rtol=1e-6, atol=0)


def adaptive_sweep(solve, fmin, fmax, tol=2e-3, ngrid=401, max_points=40, n0=5, log=print, on_step=None, seed=(),
                   grid=None, hint=None):
    """(f, S, fs, Ss, hist): the model's S (ngrid, n, n) on the grid f (ngrid linear points over fmin..fmax, or the
    given grid; a geometric grid is modelled in log f), the solved frequencies fs and their S, and one history
    entry per model: dict(solves, change, f_change, predict, f_new, converged). solve(f) -> S (n x n). The first n0
    samples spread evenly over the grid (plus the seed frequencies); each later one where the last two models
    differ most; stops when converged or at max_points solves. A non-finite S raises FloatingPointError (hint: what
    to try)."""
    fg = np.asarray(grid, float) if grid is not None else np.linspace(fmin, fmax, int(ngrid))
    use_log = grid is not None and _is_geometric(fg)
    x = (lambda f: np.log(np.asarray(f, float))) if use_log else (lambda f: np.asarray(f, float) / fg[-1])
    xg = x(fg)
    idx = sorted(set(np.round(np.linspace(0, len(fg) - 1, min(n0, len(fg)))).astype(int).tolist()))
    fs = [float(fg[i]) for i in idx] + [float(f) for f in seed if float(f) not in {float(fg[i]) for i in idx}]
    Ss = []

    def run(f):
        S = np.asarray(solve(f), complex)
        if not np.isfinite(S).all():
            raise FloatingPointError(f"the solve at {f / 1e9:.4g} GHz is not finite" + (f" ({hint})" if hint else ''))
        return S
    for f in fs:
        Ss.append(run(f))
    shape = Ss[0].shape
    hist, prev, predict = [], None, np.inf

    def fit():
        return aaa(x(fs), np.array([S.ravel() for S in Ss]))(xg)

    def add(k, model):
        """Solve grid point k; the model's prediction error there."""
        fn = float(fg[k])
        S = run(fn)
        fs.append(fn)
        Ss.append(S)
        return float(np.abs(S.ravel() - model[k]).max())
    # the first model only places the next sample: the middle of the widest gap between the first ones
    prev = fit()
    s = np.sort(np.asarray(fs))
    gap = int(np.argmax(np.diff(s)))
    k = int(np.argmin(np.where(np.isin(fg, fs), np.inf, np.abs(fg - 0.5 * (s[gap] + s[gap + 1])))))
    predict = add(k, prev)
    while True:
        # one history entry per model: its solves, its change from the model before (largest over the grid), how
        # well that one predicted the last sample, and the next sample it asks for (None when the sweep ends)
        Rg = fit()
        d = np.abs(Rg - prev).max(axis=1)
        kc = int(np.argmax(d))
        change = float(d[kc])
        converged = change <= tol and predict <= tol
        entry = dict(solves=len(fs), change=change, f_change=float(fg[kc]), predict=predict, converged=converged,
                     f_new=None)
        hist.append(entry)
        if converged or len(fs) >= max_points:
            if on_step:
                on_step(entry)
            break
        d[np.isin(fg, fs)] = -1.0                      # the next sample where the last two models differ most
        k = int(np.argmax(d))
        entry['f_new'] = float(fg[k])
        predict = add(k, Rg)
        log(f"  adaptive sweep: {len(fs)} solves, model change {change:.2e}, prediction {predict:.2e}")
        if on_step:
            on_step(entry)
        prev = Rg
    order = np.argsort(fs)
    return fg, Rg.reshape((len(fg),) + shape), np.asarray(fs)[order], np.array(Ss)[order], hist
