"""Frequency-point solves/s on one MI355X: the seeded 4-layer board (N ~ 7.8k unknowns, 6 vias, 4 via ports)
swept over 16 frequencies geomspace(1, 8 GHz), mixed precision (complex64 LU, its trailing updates accumulated on the FP64 matrix cores, FP64 GMRES
refinement, factors kept and reused across frequencies) and, beside it, the plain FP64 path (complex128 LU for
every point).

    python bench.py [--gpus 1] [--steps 16] [--warmup 2] [--no-cpu]

A step is one frequency point: Green's tables, fill, near field, vias, factor (fresh or reused), refinement and S;
the timed steps walk the sweep in order from a state with no kept factor. Prints one JSON line (value: mixed
solves/s) with the FP64 path's figures, p50 / p99 latency, the roofline of the sweep's dominant kernel (the one
with the most GPU time in an untimed attribution pass over the same sweep, every launch bracketed by HIP events;
its own launches bracketed by HIP events on their stream in the timed sweep), the rooflines of the other large
kernels and of the far fill, the LU rates, and the CPU baseline (the oracle - the reference's solve restated, its fill bit-identical to the
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


PEAK = {'hbm': (HBM_PEAK, 'GB/s', 1e9), 'fp64': (FP64_PEAK, 'TFLOP/s', 1e12), 'fp32': (FP32_PEAK, 'TFLOP/s', 1e12)}
# the roofline that bounds each kernel: HBM for the streaming ones, the matrix cores the trailing update runs on
# (FP64 for both element types of the tournament LU: its complex64 update accumulates in FP64)
BOUND = dict(fill_far='hbm', near_scatter='hbm', apply_moves='hbm', trsv='hbm', zgemm_nk='hbm', cgs2='hbm',
             laswp='hbm', gemm='fp64', near_moments='fp64', tslu='fp64', l21='fp64', trsm='fp64', panel='fp64')


def kernel_shares(rec):
    """GPU time per (kernel, dtype) over a timing record, largest first: [(kernel, dtype, total ms, launches)]."""
    out = {}
    for k, d, ms in zip(rec['kernel'], rec['dtype'], rec['ms']):
        t, c = out.get((k, d), (0.0, 0))
        out[(k, d)] = (t + ms, c + 1)
    return sorted(((k, d, t, c) for (k, d), (t, c) in out.items()), key=lambda x: -x[2])


def roofline_of(rec, kernel, dtype):
    """The roofline entry of one kernel from a timing record: achieved = the algorithmic bytes (HBM-bound kernels)
    or flops (matrix-core-bound ones) of its launches / their summed event durations."""
    sel = (rec['kernel'] == kernel) & (rec['dtype'] == dtype)
    ms = rec['ms'][sel]
    bound = BOUND[kernel]
    peak, unit, scale = PEAK[bound]
    work = rec['bytes'][sel] if bound == 'hbm' else rec['flops'][sel]
    achieved = float(work.sum() / (ms.sum() * 1e-3)) if ms.sum() > 0 else 0.0
    return dict(bound='hbm' if bound == 'hbm' else 'mfma', kernel=f"k_{kernel}<{'double2' if dtype == 'c128' else 'float2'}>",
                achieved=round(achieved / scale, 3), peak=peak / scale, unit=unit, frac=round(achieved / peak, 4),
                ms=round(float(ms.mean()), 4), calls=int(sel.sum()),
                algorithmic=float(work.mean()), algorithmic_unit='B' if bound == 'hbm' else 'flop')


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


def measured_traffic(kernel):
    """HBM bytes a dispatch of `kernel` (e.g. 'k_trsv<float2>') moved, as recorded in profiles/*_hbm_traffic.txt
    (rocprofv3 FETCH_SIZE and WRITE_SIZE of `bench.py --roofline-only`, each scaled by what the counter reads for a
    1 GiB device copy), or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, 'profiles', '*_hbm_traffic.txt'))):
        for line in open(path):
            if line.startswith(kernel + ' ') and 'MB a dispatch' in line:
                return float(line.rsplit('=', 1)[1].split('MB')[0]) * 1e6
    return None


def tslu_steps(n, leaf=256, arity=4):
    """GEPP steps on the critical path of each 64-column panel's tournament, summed over the panels of an n x n LU:
    64 a level, 1 + ceil(log_arity(leaves)) levels."""
    total = 0
    for j0 in range(0, n, 64):
        m, jb = n - j0, min(64, n - j0)
        leaves = -(-m // leaf)
        levels = 1 + (int(np.ceil(np.log(leaves) / np.log(arity) - 1e-12)) if leaves > 1 else 0)
        total += jb * levels
    return total


def roofline_only(sim, freqs):
    """The kernels whose rooflines the benchmark reports, alone, for rocprofv3 runs: the complex64 LU of the board
    at 2 GHz (k_tslu, k_gemm, ...), ten m25x_getrs calls with its four right-hand sides, ten far fills, and a 1 GiB
    device-to-device copy whose counter readings give the scale of FETCH_SIZE / WRITE_SIZE."""
    import torch
    from mom25d_amd import device as dv
    f = freqs[5]
    Z = sim.fast.fill(f)
    d = dv.diag_scales(Z)
    n = Z.shape[0]
    Z64 = torch.empty((n, n), dtype=dv.c64, device=Z.device)
    dv.scale_c64(Z, d, Z64)
    dv.timing('all')
    F = dv.getrf(Z64)
    B = torch.from_numpy(np.asarray(sim.excitation(), complex)).to(Z.device)
    for _ in range(10):
        dv.getrs(F, B)
    del F, Z64
    fo = sim.fast
    Zf = torch.empty((sim.Ne, sim.Ne), dtype=dv.c128, device=dv.device())
    for _ in range(10):
        dv.fill_far(Zf, fo.mesh, fo.tabs, fo.w)
    rec = dv.timing_records()
    dv.timing(())
    src = torch.empty(1 << 30, dtype=torch.uint8, device=Z.device)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps(dict(kernels={f"{k}/{d}": dict(ms_total=round(t, 4), launches=c)
                                   for k, d, t, c in kernel_shares(rec)}, copy_bytes=1 << 30)))


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
                    help="only the roofline kernels, for profilers: the complex64 LU of the board at 2 GHz, ten "
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
    # which kernel dominates: an untimed pass over the same sweep with every launch bracketed by events
    dv.timing('all')
    sweep(sm, freqs, a.steps)
    rec_all = dv.timing_records()
    dv.timing(())
    shares = kernel_shares(rec_all)
    dom_k, dom_d = shares[0][0], shares[0][1]
    note(f"attribution pass done: dominant kernel k_{dom_k} ({dom_d}), {shares[0][2]:.1f} ms of "
         f"{sum(x[2] for x in shares):.1f} ms")
    # the timed sweep: only the dominant kernel's launches are bracketed (the others run unobserved)
    dv.timing([dom_k])
    T, lat, Sm = sweep(sm, freqs, a.steps)
    rec = dv.timing_records()
    dv.timing(())
    note(f"timed mixed sweep: {a.steps} points in {T:.3f} s")
    value = a.steps / T
    out = dict(metric="freq-point solves/s at N=8k (mixed-prec vs FP64 plain); p50/p99 solve latency",
               value=round(value, 3), unit="solves/s", n_gpus=1, steps=a.steps, warmup=a.warmup,
               ms_per_step=round(1e3 * T / a.steps, 3), higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="complex64 LU + complex128 fill/refinement", data="synthetic",
               p50_ms=round(1e3 * float(np.percentile(lat, 50)), 3), p99_ms=round(1e3 * float(np.percentile(lat, 99)), 3))
    notes = [n for n in sm.log.notes if 'reused' in n or 'mixed' in n]
    out['reused_points'] = sum('reused' in n for n in notes[-a.steps:])
    rf = roofline_of(rec, dom_k, dom_d)
    rf['traffic'] = measured_traffic(rf['kernel'])
    rf['share_of_gpu_time'] = round(shares[0][2] / sum(x[2] for x in shares), 4)
    out['roofline'] = rf
    # the other kernels of the sweep, from the attribution pass
    tot = sum(x[2] for x in shares)
    others = {}
    for k, d, t, c in shares[:8]:
        e = roofline_of(rec_all, k, d)
        e['share_of_gpu_time'] = round(t / tot, 4)
        if k == 'tslu':
            e['bound'] = 'latency'
            e['us_per_gepp_step'] = round(1e3 * t / (c / max(1, -(-N // 64)) * tslu_steps(N)), 3)
        others[e['kernel']] = e
    out['rooflines'] = others
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
        del sd
    lu = lu_rate(sm, freqs[5])
    note("LU rates done")
    out['lu'] = {k: dict(ms=round(v['ms'], 3), tflops=round(v['tflops'], 2), peak_tflops=FP64_PEAK / 1e12,
                         frac=round(v['tflops'] * 1e12 / FP64_PEAK, 4),
                         matrix_cores='FP64 (complex64: FP64 accumulation, one rounding to FP32)' if k == 'c64'
                         else 'FP64')
                 for k, v in lu.items()}
    rff = kernel_roofline(sm, freqs[5])
    out['roofline_fill'] = dict(bound='hbm', kernel=rff['kernel'], achieved=round(rff['achieved'] / 1e9, 1),
                                peak=HBM_PEAK / 1e9, unit='GB/s', frac=round(rff['achieved'] / HBM_PEAK, 4),
                                traffic=measured_traffic('k_fill_far<double2>'), ms=round(rff['ms'], 4),
                                algorithmic_bytes=rff['bytes'])
    out['fill_pairs_per_s'] = round(sm.Ne * (sm.Ne + 1) / 2 / (rff['ms'] * 1e-3), 1)
    out['config'] = dict(workload=f"C2: seeded 4-layer board (seed 2), N={N} ({sm.Ne} RWG edges + {sm.Nv} via "
                                  f"chains), 16 frequencies geomspace(1, 8 GHz) swept in order, mixed precision",
                         N=int(N), board_seed=2, freqs_ghz=[round(f / 1e9, 4) for f in freqs], near_pairs=sm.fast.plan.n)
    if not a.no_cpu:
        note("CPU baseline (oracle, one frequency)")
        out['cpu_baseline'] = cpu_baseline(sm, freqs[5])
    print(json.dumps(out))


if __name__ == '__main__':
    main()
