"""Load generator for the serving loop (hpo_server): chemistry-substep requests arriving as a
Poisson process, as a CFD code with many partitions would issue them.

Each arrival is a request stream: a partition of log-uniform [1, 65536] cells, of the H2/air
mechanism (70 %) or GRI-3.0 (30 %), with states from the H2/air and CH4/air generators of
hpc_ode_amd.workloads. The stream is opened on the server (admitted against its HBM budget;
a refused open is retried), advances `substeps` chemistry substeps of dt ~ U[1e-5, 1e-4] s
with its state resident on the device between them, returns its final state and is closed.
The latency of a substep is the server's submit-to-result time. The requests' initial states are
drawn before the load starts, and the summary states how late the client threads started
requests against their arrival times (client_start_delay_ms_p99), so that a load generator that
falls behind shows.
"""
from __future__ import annotations

import concurrent.futures as cf
import threading
import time

import numpy as np

from . import workloads as WL

MAX_CELLS = 65536
H2_SHARE = 0.7


def request_mix(n, seed, max_cells=MAX_CELLS):
    """(cells, mechanism index 0 = H2 / 1 = GRI, dt) for n requests."""
    rng = np.random.default_rng(seed)
    cells = np.floor(np.exp(rng.uniform(0.0, np.log(max_cells + 1), n))).astype(np.int64)
    cells = np.clip(cells, 1, max_cells)
    mech = (rng.uniform(size=n) >= H2_SHARE).astype(np.int32)
    dt = rng.uniform(1e-5, 1e-4, n)
    return cells, mech, dt


def initial_states(mech_obj, which, n, seed):
    if which == 0:
        y, P, _ = WL.h2_air(mech_obj.species, mech_obj.W, n, seed=seed)
    else:
        y, P, _ = WL.gri_plain(mech_obj.species, mech_obj.W, n, seed=seed)
    return y, P


def mean_cells(max_cells=MAX_CELLS):
    """Expected cells per request of the log-uniform draw."""
    c, _, _ = request_mix(200000, 0, max_cells)
    return float(c.mean())


def run_load(server, mechs, rate_hz, duration_s, seed=0, substeps=2, workers=192, max_cells=MAX_CELLS,
             log=None, min_substeps=0):
    """Offers Poisson arrivals at rate_hz for duration_s seconds, or for as long as it takes to
    offer min_substeps substeps if that is longer; returns a summary dict."""
    rng = np.random.default_rng(seed)
    n_min = -(-min_substeps // substeps)
    gaps = rng.exponential(1.0 / rate_hz, max(int(rate_hz * duration_s * 2), 2 * n_min) + 16)
    times = np.cumsum(gaps)
    n = max(int((times < duration_s).sum()), n_min)
    times = times[:n]
    duration_s = max(duration_s, float(times[-1])) if n else duration_s
    cells, mech, dt = request_mix(n, seed + 1, max_cells)
    # every request's initial state is drawn before the clock starts, so the client threads do
    # nothing at arrival but call the server (drawing them on arrival held the interpreter long
    # enough to start 1 % of the requests more than a second late)
    states = [initial_states(mechs[mech[i]], mech[i], int(cells[i]), seed * 1000003 + i) for i in range(n)]
    lat, queue, bsz, breq, late = [], [], [], [], []
    failed = [0]
    integrated = [0]
    refused = [0]
    lock = threading.Lock()

    def one(i):
        with lock:
            late.append(time.perf_counter() - t0 - times[i])
        m = mechs[mech[i]]
        y0, P = states[i]
        sid = None
        while sid is None:
            sid = server.open_stream(int(mech[i]), y0, P)
            if sid is None:
                with lock:
                    refused[0] += 1
                time.sleep(0.001)
        try:
            for s in range(substeps):
                last = s == substeps - 1
                rid = server.submit(sid, float(dt[i]), want_state=last, n=int(cells[i]), n_state=m.n_state)
                _, st, info = server.wait(rid)
                with lock:
                    lat.append(info["latency_ms"])
                    queue.append(info["queue_ms"])
                    bsz.append(info["batch_cells"])
                    breq.append(info["batch_requests"])
                    failed[0] += info["failed_cells"]
                    integrated[0] += int(cells[i])
        finally:
            server.close_stream(sid)

    t0 = time.perf_counter()
    futs = []
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for i in range(n):
            delay = times[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            futs.append(ex.submit(one, i))
        for f in futs:
            f.result()
    wall = time.perf_counter() - t0
    lat = np.asarray(lat)
    out = {
        "offered_requests_per_s": rate_hz,
        "offered_cells_per_s": float(cells.sum() * substeps / duration_s),
        "requests": int(n), "substeps_per_request": substeps,
        "h2_share": float((mech == 0).mean()) if n else None,
        "substeps_completed": int(lat.size),
        "latency_ms_p50": float(np.percentile(lat, 50)) if lat.size else None,
        "latency_ms_p99": float(np.percentile(lat, 99)) if lat.size else None,
        "latency_ms_max": float(lat.max()) if lat.size else None,
        "queue_ms_p50": float(np.percentile(queue, 50)) if queue else None,
        "client_start_delay_ms_p99": float(np.percentile(late, 99) * 1e3) if late else None,
        "mean_batch_cells": float(np.mean(bsz)) if bsz else None,
        "mean_batch_requests": float(np.mean(breq)) if breq else None,
        "cells_per_s": integrated[0] / wall,
        "wall_s": wall,
        "failed_cells": failed[0],
        "refused_opens": refused[0],
    }
    if log:
        log(f"[serving] {rate_hz:.1f} req/s: p50 {out['latency_ms_p50']} ms p99 {out['latency_ms_p99']} ms, "
            f"{out['cells_per_s']:.0f} cells/s, batch {out['mean_batch_requests']} req")
    return out


def capacity(server, mechs, seed=7, n=48, max_cells=MAX_CELLS):
    """Cells per second the server sustains when n requests of the mix are all queued at once."""
    cells, mech, dt = request_mix(n, seed, max_cells)
    sids = []
    for i in range(n):
        m = mechs[mech[i]]
        y0, P = initial_states(m, mech[i], int(cells[i]), seed + i)
        sids.append(server.open_stream(int(mech[i]), y0, P))
    t0 = time.perf_counter()
    rids = [server.submit(sids[i], float(dt[i]), want_state=False) for i in range(n)]
    for r in rids:
        server.wait(r)
    wall = time.perf_counter() - t0
    for s in sids:
        server.close_stream(s)
    return float(cells.sum() / wall)
