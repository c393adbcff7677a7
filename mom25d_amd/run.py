"""S-parameters of a Simulation over a band at the output reference impedances (src/mom25d/run.py:14-84).

    f, S, info = solve_band(sim, 1e9, 6e9, linear=101, zref=50)

Every frequency point is one GPU solve (tables, fill, near field, vias, factor or kept factor, refinement);
S is renormalised from each port's own impedance to zref. The reference's default adaptive sweep (a rational fit
through emsdev.afs, not available offline) is not built: ask for `linear` points.
"""
import time

import numpy as np


def renorm(S, zc, zref):
    """S renormalised from reference impedances zc to zref (scalars or per port, real) (run.py:14-23)."""
    zc, n = np.asarray(zc, float), S.shape[0]
    zc = np.broadcast_to(zc, (n,))
    zref = np.broadcast_to(np.asarray(zref, float), (n,))
    if np.allclose(zc, zref):
        return S
    I = np.eye(n)
    sq = np.diag(np.sqrt(zc))
    Z = sq @ (I + S) @ np.linalg.inv(I - S) @ sq
    isq = np.diag(1 / np.sqrt(zref))
    zn = isq @ Z @ isq
    return (zn - I) @ np.linalg.inv(zn + I)


def out_refs(sim, z0):
    """Output reference per port: z0 (a number, or {port name: ohms}) (run.py:26-33; lumped ports have no modes)."""
    if isinstance(z0, dict):
        return np.array([float(z0.get(p.name, 50.0)) for p in sim.ports])
    return np.full(len(sim.ports), float(z0))


def _zc(sim, f):
    return np.array([p.zc_fn(f) if getattr(p, 'zc_fn', None) else p.zc for p in sim.ports], float)


def solve_band(sim, fmin, fmax, points=401, zref=50.0, linear=None, log=print, progress=None, keep_at=(),
               freq=None):
    """(f, S (nf, n, n), info) of sim over fmin..fmax (Hz) (run.py:41-84): `linear` equally spaced points, or the
    frequencies `freq` when given. info: solves, seconds, zref, the solved frequencies, port names. keep_at:
    frequencies whose currents are kept in sim.kept. progress(solves, total, f) after each solve."""
    if freq is None and not linear:
        raise ValueError("solve_band: the adaptive sweep is not available in this build; pass linear=N (or freq=)")
    t0, zr, done = time.time(), out_refs(sim, zref), []
    keep_at = {float(f) for f in keep_at}
    sim.kept = {}
    fs = np.asarray(freq, float) if freq is not None else np.linspace(fmin, fmax, int(linear))
    out = []
    for fr in fs:
        log(f"  solving {fr / 1e9:.4f} GHz")
        keep = float(fr) in keep_at
        out.append(renorm(sim.sparams(fr, keep=keep), _zc(sim, fr), zr))
        if keep:
            sim.kept[float(fr)] = sim.last['I']
        done.append(float(fr))
        if progress:
            progress(len(done), len(fs), fr)
    return fs, np.array(out), dict(solves=len(fs), seconds=time.time() - t0, zref=zr.tolist(), solved=done,
                                   names=[p.name for p in sim.ports], afs=[], afs_tol=None)
