"""Frequency-point solves/s on one MI355X: the seeded 4-layer board (N ~ 7.8k unknowns, 6 vias, 4 via ports)
swept over 16 frequencies geomspace(1, 8 GHz), mixed precision (complex64 LU on the FP32 matrix cores, FP64 GMRES
refinement, factors kept and reused across frequencies) and, beside it, the plain FP64 path (complex128 LU for
every point).

    python bench.py [--gpus 1] [--steps 16] [--warmup 2] [--no-cpu]

A step is one frequency point: Green's tables, fill, near field, vias, factor (fresh or reused), refinement and S;
the timed steps walk the sweep in order from a state with no kept factor. Prints one JSON line (value: mixed
solves/s) with the FP64 path's figures, p50 / p99 latency, the roofline of the dominant single kernel (HIP events
around its launches, on the stream it runs on) and the CPU baseline (the oracle - the reference's solve restated,
its fill bit-identical to the reference's compiled kernels - with SciPy's LAPACK, on this host's cores, one
frequency of the same board).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
FP32_PEAK = 157.3e12
FP64_PEAK = 78.6e12


def build_sim(precision):
    from mom25d_amd import boards
    from mom25d_amd.sim import Simulation
    st, sh, po, vi, kw = boards.seeded_board()
    return Simulation(st, sh, po, vi, precision=precision, verbose=False, **kw)


def sweep(sim, freqs, steps):
    import torch
    fo = sim.fast
    fo._kept, fo._reach, fo._t_iter, fo._mixed_failed = [], np.log(2.0), None, 0.0
    lat, S = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        f = freqs[i % len(freqs)]
        t = time.perf_counter()
        S.append(sim.sparams(f))
        lat.append(time.perf_counter() - t)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, np.array(lat), S


def kernel_roofline(sim, f, reps=5):
    """The far fill (one launch, m25x_fill_far): HIP events on the current stream around each launch; algorithmic
    bytes = its Z stores, 16 B per entry of the Ne x Ne edge block."""
    import torch
    from mom25d_amd import device as dv
    fo = sim.fast
    fo.setup(f)
    Ne = sim.Ne
    Z = torch.empty((Ne, Ne), dtype=dv.c128, device=dv.device())
    dv.fill_far(Z, fo.mesh, fo.tabs, fo.w)
    s = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record(s)
        dv.fill_far(Z, fo.mesh, fo.tabs, fo.w)
        b.record(s)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    traffic = 16.0 * Ne * Ne
    achieved = traffic / (ms * 1e-3)
    return dict(kernel='m25x_fill_far (k_fill_far<double2>)', ms=ms, bytes=traffic, achieved=achieved)


def lu_rate(sim, f):
    """Complex LU rates on this board's Z (8 N^3 / 3 flops): complex64 and complex128, events around getrf."""
    import torch
    from mom25d_amd import device as dv
    Z = sim.fast.fill(f)
    d = dv.diag_scales(Z)
    n = Z.shape[0]
    Z64 = torch.empty((n, n), dtype=dv.c64, device=Z.device)
    dv.scale_c64(Z, d, Z64)
    out = {}
    s = torch.cuda.current_stream()
    for name, A in (('c64', Z64), ('c128', Z)):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        W = A.clone()
        a.record(s)
        dv.getrf(W)
        b.record(s)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        out[name] = dict(ms=ms, tflops=8.0 * n ** 3 / 3 / (ms * 1e-3) / 1e12)
    return out


def cpu_baseline(sim, f):
    """One frequency of the same board through the oracle on this host's cores (FP64 path: tables, fill, near
    field, vias, zgetrf / zgetrs): solves/s."""
    sys.path.insert(0, ROOT)
    from oracle.solve import Solver
    from tests.cases import oracle_board
    cores = len(os.sched_getaffinity(0))
    b = oracle_board(sim)
    t0 = time.perf_counter()
    osv = Solver(b)
    t_geom = time.perf_counter() - t0
    t0 = time.perf_counter()
    osv.solve(f)
    t = time.perf_counter() - t0
    return dict(value=1.0 / t, unit='solves/s', cores=cores, kind='port',
                sample=f"one frequency ({f / 1e9:.3g} GHz) of the same N={sim.Ne + sim.Nv} board, FP64: tables + fill + "
                       f"near field + vias + LAPACK zgetrf/zgetrs (SciPy), {t:.2f} s (+{t_geom:.2f} s of "
                       f"per-board geometry not counted)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=16)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-double', action='store_true')
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: frequency points and boards are independent "
                                   "and would shard across GPUs with no exchange (see DESIGN.md); --gpus 1"}))
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    freqs = np.geomspace(1e9, 8e9, 16)
    sm = build_sim('mixed')
    N = sm.Ne + sm.Nv
    for i in range(a.warmup):
        sm.sparams(freqs[i % len(freqs)])
    T, lat, Sm = sweep(sm, freqs, a.steps)
    value = a.steps / T
    out = dict(metric="freq-point solves/s at N=8k (mixed-precision vs FP64 plain); p50/p99 solve latency",
               value=round(value, 3), unit="solves/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=round(1e3 * T / a.steps, 3), higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="complex64 LU + complex128 fill/refinement", data="synthetic",
               p50_ms=round(1e3 * float(np.percentile(lat, 50)), 3), p99_ms=round(1e3 * float(np.percentile(lat, 99)), 3))
    notes = [n for n in sm.log.notes if 'reused' in n or 'mixed' in n]
    out['reused_points'] = sum('reused' in n for n in notes[-a.steps:])
    if not a.no_double:
        sd = build_sim('double')
        for i in range(min(a.warmup, 1)):
            sd.sparams(freqs[i])
        nd = min(a.steps, 8)
        Td, latd, Sd = sweep(sd, freqs, nd)
        out['fp64_plain'] = dict(value=round(nd / Td, 3), steps=nd, ms_per_step=round(1e3 * Td / nd, 3),
                                 p50_ms=round(1e3 * float(np.percentile(latd, 50)), 3),
                                 p99_ms=round(1e3 * float(np.percentile(latd, 99)), 3))
        out['mixed_over_fp64'] = round(value / (nd / Td), 3)
        out['max_dS_mixed_vs_fp64'] = float(max(np.abs(Sm[i] - Sd[i]).max() for i in range(nd)))
    lu = lu_rate(sm, freqs[5])
    out['lu'] = {k: dict(ms=round(v['ms'], 3), tflops=round(v['tflops'], 2),
                         frac=round(v['tflops'] * 1e12 / (FP32_PEAK if k == 'c64' else FP64_PEAK), 4))
                 for k, v in lu.items()}
    rf = kernel_roofline(sm, freqs[5])
    out['roofline'] = dict(bound='hbm', kernel=rf['kernel'], achieved=round(rf['achieved'] / 1e9, 1),
                           peak=HBM_PEAK / 1e9, unit='GB/s', frac=round(rf['achieved'] / HBM_PEAK, 4),
                           traffic=None, ms=round(rf['ms'], 4), algorithmic_bytes=rf['bytes'])
    out['fill_pairs_per_s'] = round(sm.Ne * (sm.Ne + 1) / 2 / (rf['ms'] * 1e-3), 1)
    out['config'] = dict(workload=f"C2: seeded 4-layer board (seed 2), N={N} ({sm.Ne} RWG edges + {sm.Nv} via "
                                  f"chains), 16 frequencies geomspace(1, 8 GHz) swept in order, mixed precision",
                         N=int(N), board_seed=2, freqs_ghz=[round(f / 1e9, 4) for f in freqs], near_pairs=sm.fast.plan.n)
    if not a.no_cpu:
        out['cpu_baseline'] = cpu_baseline(sm, freqs[5])
    print(json.dumps(out))


if __name__ == '__main__':
    main()
