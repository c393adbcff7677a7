"""S-parameters of a Simulation over a band at the output reference impedances (src/mom25d/run.py:14-84).

    f, S, info = solve_band(sim, 1e9, 6e9, points=401, zref=50)          # adaptive: a rational model, few solves
    f, S, info = solve_band(sim, 1e9, 6e9, linear=101, zref=50)          # 101 solves

Every frequency point solved is one GPU solve (tables, fill, near field, vias, factor or kept factor, refinement);
S is renormalised from each port's own impedance to zref. The adaptive sweep is mom25d_amd.afs (the build's own
AAA sweep in place of emsdev.afs).
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
    """Output reference per port: z0 (a number, {port name: ohms} or one value a port) (run.py:26-33; lumped ports
    have no modes)."""
    if isinstance(z0, dict):
        return np.array([float(z0.get(p.name, 50.0)) for p in sim.ports])
    if np.ndim(z0):
        z = np.asarray(z0, float)
        if z.shape != (len(sim.ports),):
            raise ValueError(f"zref: {len(z)} values for {len(sim.ports)} ports")
        return z
    return np.full(len(sim.ports), float(z0))


def _zc(sim, f):
    return np.array([p.zc_fn(f) if getattr(p, 'zc_fn', None) else p.zc for p in sim.ports], float)


def solve_band(sim, fmin, fmax, points=401, zref=50.0, linear=None, afs_tol=2e-3, max_solves=40, log=print,
               keep_mid=False, progress=None, on_step=None, keep_at=(), freq=None):
    """(f, S (nf, n, n), info) of sim over fmin..fmax (Hz) (run.py:41-84). Adaptive by default: a rational model
    through as few solves as afs_tol (max |dS| between successive models, and of the model's prediction of each new
    sample) allows, at most max_solves, evaluated on `points` equally spaced frequencies, or on `freq` when given
    (a geometric `freq` is modelled in log f); linear=N solves N equally spaced frequencies instead. keep_mid: one
    more solve at mid-band keeping its currents (sim.last). keep_at: frequencies the sweep solves among its samples
    and keeps the currents of (sim.kept). progress(solves, at most, f) after each solve; on_step(entry) after each
    model. info: solves, seconds, zref, the solved frequencies, port names, the adaptive sweep's history and tol."""
    t0, zr, done = time.time(), out_refs(sim, zref), []
    keep_at = {float(f) for f in keep_at}
    sim.kept = {}

    def solve(fr):
        log(f"  solving {fr / 1e9:.4f} GHz")
        keep = float(fr) in keep_at
        S = renorm(sim.sparams(fr, keep=keep), _zc(sim, fr), zr)
        if keep:
            sim.kept[float(fr)] = sim.last['I']
        done.append(float(fr))
        if progress:
            progress(len(done), int(linear) if linear else max_solves, fr)
        return S
    hist = []
    if linear:
        f = fs = np.linspace(fmin, fmax, int(linear))
        S = np.array([solve(fi) for fi in f])
    else:
        from .afs import adaptive_sweep
        f, S, fs, _, hist = adaptive_sweep(solve, fmin, fmax, tol=afs_tol, ngrid=points, max_points=max_solves,
                                           log=log, on_step=on_step, seed=sorted(keep_at), grid=freq)
        if hist and not hist[-1]['converged']:
            sim.log.note(f"WARNING adaptive sweep: stopped at its {max_solves} solves before its models agreed (last "
                         f"change {hist[-1]['change']:.1e}, tolerance {afs_tol:g}): more solves or a narrower band")
    if keep_mid:
        sim.sparams(fs[len(fs) // 2], keep=True)
    return f, S, dict(solves=len(fs), seconds=time.time() - t0, zref=zr.tolist(), solved=list(map(float, fs)),
                      names=[p.name for p in sim.ports], afs=hist, afs_tol=None if linear else afs_tol)
