"""Frequency-point solves/s on one MI355X: the seeded 4-layer board (N ~ 7.8k unknowns, 6 vias, 4 via ports)
swept over 16 frequencies geomspace(1, 8 GHz), mixed precision (complex64 LU on the FP32 matrix cores, FP64 GMRES
refinement, factors kept and reused across frequencies) and, beside it, the plain FP64 path (complex128 LU for
every point).

    python bench.py [--gpus 1] [--steps 16] [--warmup 2] [--no-cpu]

A step is one frequency point: Green's tables, fill, near field, vias, factor (fresh or reused), refinement and S;
the timed steps walk the sweep in order from a state with no kept factor. Prints one JSON line (value: mixed
solves/s) with the FP64 path's figures, p50 / p99 latency, the roofline of the sweep's dominant kernels - the
triangular solves of the refinement, HIP events on their stream around every call in the timed sweep - and of the
far fill, and the CPU baseline (the oracle - the reference's solve restated, its fill bit-identical to the
reference's compiled kernels - with SciPy's LAPACK, on this host's cores, one frequency of the same board).
--roofline-only runs just those kernels (for rocprofv3).
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


def solve_roofline(n, es):
    """The triangular solves (k_trsv: U^T, then L^T, in each m25x_getrs call), the sweep's dominant kernel: the
    library's HIP events on the launching stream around every launch recorded since dv.trsv_timing(True).
    Algorithmic bytes per launch = its triangle of the factor (n (n +- 1) / 2 elements) and the right-hand sides
    read and written once."""
    from mom25d_amd import device as dv
    ms, by = dv.trsv_times()
    # the launches on a factor of element size es (the sweep's FP64 fall-backs read twice the bytes)
    lim = (n * (n + 1) / 2 + 16.0 * n) * 8
    sel = by <= lim if es == 8 else by > lim
    ms, by = ms[sel], by[sel]
    avg = float(ms.mean())
    return dict(kernel=f"k_trsv<{'float2' if es == 8 else 'double2'}> (m25x_getrs, N={n}, up to 4 right-hand sides)",
                ms=avg, calls=int(sel.sum()), bytes=float(by.mean()), achieved=float(by.mean() / (avg * 1e-3)))


def measured_traffic(kernel):
    """HBM bytes a dispatch of `kernel` moved, as recorded in profiles/trsv_fill_board_n7799_hbm_traffic.txt
    (rocprofv3 FETCH_SIZE and WRITE_SIZE of `bench.py --roofline-only`, scaled by a 1 GiB copy), or None."""
    try:
        for line in open(os.path.join(ROOT, 'profiles', 'trsv_fill_board_n7799_hbm_traffic.txt')):
            if line.startswith(kernel + ' ') and 'MB a dispatch' in line:
                return float(line.rsplit('=', 1)[1].split('MB')[0]) * 1e6
    except OSError:
        pass
    return None


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


def roofline_only(sim, freqs):
    """The kernels whose rooflines the benchmark reports, alone, for rocprofv3 runs: ten m25x_getrs calls (four
    right-hand sides) with the complex64 factor of the board at 2 GHz, ten far fills, and a 1 GiB device-to-device
    copy whose counter readings give the scale of FETCH_SIZE / WRITE_SIZE."""
    import torch
    from mom25d_amd import device as dv
    f = freqs[5]
    Z = sim.fast.fill(f)
    d = dv.diag_scales(Z)
    n = Z.shape[0]
    Z64 = torch.empty((n, n), dtype=dv.c64, device=Z.device)
    dv.scale_c64(Z, d, Z64)
    F = dv.getrf(Z64)
    B = torch.from_numpy(np.asarray(sim.excitation(), complex)).to(Z.device)
    dv.trsv_timing(True)
    for _ in range(10):
        dv.getrs(F, B)
    dv.trsv_timing(False)
    sr = solve_roofline(n, 8)
    del F, Z64
    rf = kernel_roofline(sim, f, reps=10)
    src = torch.empty(1 << 30, dtype=torch.uint8, device=Z.device)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps(dict(getrs_ms=sr['ms'], getrs_bytes=sr['bytes'], fill_ms=rf['ms'], fill_bytes=rf['bytes'],
                          copy_bytes=1 << 30)))


def note(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=16)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-double', action='store_true')
    ap.add_argument('--roofline-only', action='store_true',
                    help="only the roofline kernels, for profilers: the complex64 factor of the board at 2 GHz, ten "
                         "solves with 4 right-hand sides, ten far fills, and a 1 GiB device copy (the counters' scale)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU: frequency points and boards are independent "
                                   "and would shard across GPUs with no exchange; --gpus 1"}))
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    freqs = np.geomspace(1e9, 8e9, 16)
    sm = build_sim('mixed')
    N = sm.Ne + sm.Nv
    note(f"board built, N = {N}")
    if a.roofline_only:
        roofline_only(sm, freqs)
        return
    from mom25d_amd import device as dv
    for i in range(a.warmup):
        sm.sparams(freqs[i % len(freqs)])
    note(f"{a.warmup} warm-up points done")
    dv.trsv_timing(True)
    T, lat, Sm = sweep(sm, freqs, a.steps)
    dv.trsv_timing(False)
    note(f"timed mixed sweep: {a.steps} points in {T:.3f} s")
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
        note(f"FP64 sweep: {nd} points in {Td:.3f} s")
        out['fp64_plain'] = dict(value=round(nd / Td, 3), steps=nd, ms_per_step=round(1e3 * Td / nd, 3),
                                 p50_ms=round(1e3 * float(np.percentile(latd, 50)), 3),
                                 p99_ms=round(1e3 * float(np.percentile(latd, 99)), 3))
        out['mixed_over_fp64'] = round(value / (nd / Td), 3)
        out['max_dS_mixed_vs_fp64'] = float(max(np.abs(Sm[i] - Sd[i]).max() for i in range(nd)))
    lu = lu_rate(sm, freqs[5])
    note("LU rates done")
    out['lu'] = {k: dict(ms=round(v['ms'], 3), tflops=round(v['tflops'], 2),
                         frac=round(v['tflops'] * 1e12 / (FP32_PEAK if k == 'c64' else FP64_PEAK), 4))
                 for k, v in lu.items()}
    sr = solve_roofline(N, 8)
    out['roofline'] = dict(bound='hbm', kernel=sr['kernel'], achieved=round(sr['achieved'] / 1e9, 1),
                           peak=HBM_PEAK / 1e9, unit='GB/s', frac=round(sr['achieved'] / HBM_PEAK, 4),
                           traffic=measured_traffic('k_trsv<float2>'), ms=round(sr['ms'], 4), calls=sr['calls'],
                           algorithmic_bytes=sr['bytes'])
    rf = kernel_roofline(sm, freqs[5])
    out['roofline_fill'] = dict(bound='hbm', kernel=rf['kernel'], achieved=round(rf['achieved'] / 1e9, 1),
                                peak=HBM_PEAK / 1e9, unit='GB/s', frac=round(rf['achieved'] / HBM_PEAK, 4),
                                traffic=measured_traffic('k_fill_far<double2>'), ms=round(rf['ms'], 4),
                                algorithmic_bytes=rf['bytes'])
    out['fill_pairs_per_s'] = round(sm.Ne * (sm.Ne + 1) / 2 / (rf['ms'] * 1e-3), 1)
    out['config'] = dict(workload=f"C2: seeded 4-layer board (seed 2), N={N} ({sm.Ne} RWG edges + {sm.Nv} via "
                                  f"chains), 16 frequencies geomspace(1, 8 GHz) swept in order, mixed precision",
                         N=int(N), board_seed=2, freqs_ghz=[round(f / 1e9, 4) for f in freqs], near_pairs=sm.fast.plan.n)
    if not a.no_cpu:
        note("CPU baseline (oracle, one frequency)")
        out['cpu_baseline'] = cpu_baseline(sm, freqs[5])
    print(json.dumps(out))


if __name__ == '__main__':
    main()
