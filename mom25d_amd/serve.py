"""A request loop that serves frequency-point solves of many resident boards on one GPU.

The reference has no server: a sweep asks Simulation.sparams one frequency at a time (src/mom25d/run.py:55-64) and
its memory plan decides, per board, whether a solve fits (src/mom25d/sim.py:85-151, src/mom25d/gpu.py:37-41). Here:

  resident   every board's mesh, RWG geometry, near-pair plan and analytic near moments stay on the device
             (its Simulation); a request is (board, frequency)
  admission  a request holds N^2 (16 + 8) bytes while it is solved (its FP64 Z and its complex64 factor) plus its
             solve workspaces; requests are admitted in arrival order while the bytes of everything in flight stay
             within the budget (default 0.85 x the device's free memory when the server starts); a request larger
             than the budget alone is refused
  batching   the admitted requests form a batch (mom25d_amd.batch.SweepBatch: one batched LU, one GMRES per size
             group); a batch's first stage (fill + factor) runs on one stream and its second (refinement + S) on
             another, so the next batch - the requests that arrived while this one was on the GPU, admitted as
             memory frees up - fills and factors while this one refines: their work overlaps on the device
  results    per request: S, arrival, start and finish times, latency, its batch

A request's answer does not depend on what else is in its batch beyond rounding: its fill, its LU (the batched LU
is bit-identical to the one-matrix LU) and its GMRES columns (each column's Krylov space and stopping are its own)
are computed as if alone; only the orthogonalisation's sums may be grouped differently.
"""
import time
from collections import deque

import numpy as np
import torch

from .batch import SweepBatch


def request_bytes(sim):
    """Device bytes one request holds while it is solved: the FP64 Z (16 N^2), the complex64 factor (8 N^2), its LU
    and solve workspaces and right-hand sides (bounded by 2 KiB per unknown), and its near pairs' moments (128 B a
    pair, held from the batch's preparation to its fill)."""
    n = sim.Ne + sim.Nv
    plan = getattr(getattr(sim, 'fast', None), 'plan', None)
    return 24.0 * n * n + 2048.0 * n + 128.0 * (plan.n if plan is not None else 0)


class Server:
    """Serves (board, frequency) requests over the resident boards {name: Simulation}."""

    def __init__(self, boards, budget=None, max_batch=64):
        self.boards = dict(boards)
        if budget is None:
            free, _ = torch.cuda.mem_get_info()
            budget = 0.85 * free
        self.budget, self.max_batch = float(budget), int(max_batch)
        self.streams = (torch.cuda.Stream(), torch.cuda.Stream())

    def _admit(self, pending, in_flight):
        """Requests from the head of the queue while the bytes in flight stay within the budget."""
        batch, used = [], in_flight
        while pending and len(batch) < self.max_batch:
            r = pending[0]
            need = request_bytes(self.boards[r['board']])
            if need > self.budget:
                pending.popleft()
                r['error'] = f"needs {need / 1e9:.1f} GB, more than the budget of {self.budget / 1e9:.1f} GB"
                continue
            if used + need > self.budget:
                break
            batch.append(pending.popleft())
            used += need
        return batch, used

    def serve(self, requests, pipeline=True):
        """Runs a stream of requests [(arrival seconds, board name, frequency Hz)] (arrival times from the start of the
        call, in any order) to completion; returns one record per request, in the order given: dict(board, f,
        arrival, start, finish, latency, batch, S) (or 'error' for a refused request)."""
        recs = [dict(board=b, f=float(f), arrival=float(t), idx=i) for i, (t, b, f) in enumerate(requests)]
        for r in recs:
            if r['board'] not in self.boards:
                raise KeyError(f"no resident board {r['board']!r}")
        waiting = deque(sorted(recs, key=lambda r: (r['arrival'], r['idx'])))
        pending = deque()
        sa, sb = self.streams
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        running = None                       # (SweepBatch, its requests, its bytes, the event its stage one records)
        nbatch = 0
        while waiting or pending or running is not None:
            while waiting and waiting[0]['arrival'] <= now():
                pending.append(waiting.popleft())
            if running is None and not pending:
                if waiting:                  # idle: until the next arrival
                    time.sleep(max(0.0, min(0.01, waiting[0]['arrival'] - now())))
                continue
            # admit the next batch within what the running one leaves free, and start its first stage
            nxt = None
            in_flight = running[2] if running is not None else 0.0
            if pending and (pipeline or running is None):
                reqs, used = self._admit(pending, in_flight)
                if reqs:
                    b = SweepBatch([(self.boards[r['board']], r['f']) for r in reqs])
                    t_start = now()
                    for r in reqs:
                        r['start'], r['batch'] = t_start, nbatch
                    with torch.cuda.stream(sa):
                        b.fill_factor()
                        ev = torch.cuda.Event()
                        ev.record(sa)
                    nxt = (b, reqs, used - in_flight, ev)
                    nbatch += 1
            # finish the running batch's second stage (it overlaps the next batch's first stage on the device)
            if running is not None:
                self._finish(running, sb, now)
                running = None
            if nxt is not None:
                if not pipeline:
                    self._finish(nxt, sb, now)
                else:
                    running = nxt
        for r in recs:
            r.pop('idx', None)
        return recs

    def _finish(self, run, stream, now):
        b, reqs, _, ev = run
        stream.wait_event(ev)
        for t in b.tensors():
            t.record_stream(stream)
        with torch.cuda.stream(stream):
            S = b.refine()
        t = now()
        for r, s in zip(reqs, S):
            r['S'], r['finish'] = s, t
            r['latency'] = t - r['arrival']


def summarize(recs):
    """p50 / p99 latency (s), solves/s over the stream (completed requests / time to the last finish), mean batch."""
    done = [r for r in recs if 'S' in r]
    if not done:
        return dict(completed=0)
    lat = np.array([r['latency'] for r in done])
    span = max(r['finish'] for r in done) - min(r['arrival'] for r in done)
    sizes = np.bincount([r['batch'] for r in done])
    return dict(completed=len(done), refused=len(recs) - len(done), p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), solves_per_s=len(done) / max(span, 1e-9),
                mean_batch=float(sizes[sizes > 0].mean()), batches=int((sizes > 0).sum()))


def adaptive_order(fmin, fmax, n):
    """n frequencies in the order an adaptive sweep proposes them: the band's ends, then midpoints of the widest
    gaps (breadth first), as a model refined where it is least constrained would ask for them."""
    out, gaps = [fmin, fmax], [(fmin, fmax)]
    while len(out) < n:
        nxt = []
        for a, b in gaps:
            m = 0.5 * (a + b)
            out.append(m)
            nxt += [(a, m), (m, b)]
            if len(out) >= n:
                break
        gaps = nxt
    return out[:n]
