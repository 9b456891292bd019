"""Serving loop (hpo_server): continuous batching of substep requests over resident streams.

A request's result is the same alone and in a mix (bit-identical: each cell is integrated by one
wave with the same instruction sequence, whatever it is batched with); successive substeps of a
stream continue from its resident state; opens are admitted against the HBM budget; a stream
has one substep in flight; requests that arrive together share launches."""
import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import serving as SV
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mechs(gri_path, h2_path):
    return [H.Mechanism(h2_path), H.Mechanism(gri_path)]


def _state(m, which, n, seed):
    return SV.initial_states(m, which, n, seed)


@pytest.mark.parametrize("cu_share", [0, 30])
def test_result_same_alone_and_in_mix(mechs, cu_share):
    """With cu_share 30 every launch takes at most 30 % of the device's workgroup slots, so
    the mix runs as concurrent launches of smaller grids."""
    srv = H.Server(mechs, n_slots=3, cu_share_pct=cu_share)
    try:
        gri = mechs[1]
        y0, P = _state(gri, 1, 100, 5)
        sid = srv.open_stream(1, y0, P)
        ya, sa, ia = srv.wait(srv.submit(sid, 5e-5, n=100, n_state=gri.n_state))
        srv.close_stream(sid)
        assert ia["batch_requests"] == 1 and (sa == 0).all()
        # the same cells among other requests of both mechanisms, all submitted together
        others = []
        for k, (which, n) in enumerate([(0, 3000), (1, 700), (0, 1), (1, 4096), (0, 20000), (1, 33)]):
            yk, Pk = _state(mechs[which], which, n, 100 + k)
            others.append((which, n, srv.open_stream(which, yk, Pk)))
        sid = srv.open_stream(1, y0, P)
        rids = [srv.submit(s, 7e-5 if w else 3e-5, n=n, n_state=mechs[w].n_state) for w, n, s in others[:3]]
        rid = srv.submit(sid, 5e-5, n=100, n_state=gri.n_state)
        rids += [srv.submit(s, 2e-5, n=n, n_state=mechs[w].n_state) for w, n, s in others[3:]]
        yb, sb, ib = srv.wait(rid)
        for r in rids:
            srv.wait(r)
        assert ib["batch_requests"] > 1 or ib["batch_cells"] > 100
        assert np.array_equal(ya, yb) and np.array_equal(sa, sb)
        # and the same as the one-shot integration entry point with the server's default
        # options (FP32 Newton matrix with the FP64 fallback)
        yc, sc, _ = gri.integrate(y0, P, np.full(100, 5e-5), lu_precision=1)
        assert np.array_equal(ya, yc)
    finally:
        srv.close()


def test_substeps_continue_from_resident_state(mechs):
    srv = H.Server(mechs)
    try:
        h2 = mechs[0]
        y0, P = _state(h2, 0, 257, 9)
        sid = srv.open_stream(0, y0, P)
        srv.wait(srv.submit(sid, 4e-5, want_state=False))
        y2, s2, _ = srv.wait(srv.submit(sid, 6e-5, n=257, n_state=h2.n_state))
        assert np.array_equal(srv.stream_state(sid, 257, h2.n_state), y2)
        srv.close_stream(sid)
        y1, _, _ = h2.integrate(y0, P, np.full(257, 4e-5), lu_precision=1)
        yr, _, _ = h2.integrate(y1, P, np.full(257, 6e-5), lu_precision=1)
        assert np.array_equal(y2, yr)
    finally:
        srv.close()


def test_admission_against_hbm_budget(mechs):
    """A stream is charged the pool block that holds it: its rows, pressures and statuses
    (53 x 8 + 8 + 4 bytes per GRI cell, 256-byte aligned parts) rounded up to a power of two,
    here 256 KiB for 600 cells. A budget of 400,000 bytes admits one such stream, not two."""
    gri = mechs[1]
    srv = H.Server(mechs, hbm_budget_bytes=400_000)
    try:
        y, P = _state(gri, 1, 600, 1)
        a = srv.open_stream(1, y, P)
        assert a is not None
        assert srv.open_stream(1, y, P) is None
        used, budget = srv.hbm()
        assert used == 1 << 18 and budget == 400_000
        srv.close_stream(a)
        assert srv.hbm()[0] == 0
        b = srv.open_stream(1, y, P)          # reuses the pooled block
        assert b is not None
        assert np.array_equal(srv.stream_state(b, 600, gri.n_state), y)
        srv.close_stream(b)
    finally:
        srv.close()


def test_wait_after_close_of_a_completed_stream(mechs):
    """A stream may be closed once its substep has completed and before that request is waited
    for; the wait still reports the request (the request keeps the stream alive)."""
    srv = H.Server(mechs)
    try:
        h2 = mechs[0]
        y, P = _state(h2, 0, 300, 4)
        sid = srv.open_stream(0, y, P)
        rid = srv.submit(sid, 2e-5, want_state=False)
        import time
        for _ in range(20000):
            try:
                srv.close_stream(sid)
                break
            except H.HpoError:
                time.sleep(0.001)
        else:
            raise AssertionError("the substep did not complete")
        _, _, info = srv.wait(rid)
        assert info["ncells"] == 300 and info["failed_cells"] == 0
        # the freed block serves the next stream
        sid2 = srv.open_stream(0, y, P)
        assert np.array_equal(srv.stream_state(sid2, 300, h2.n_state), y)
        srv.close_stream(sid2)
    finally:
        srv.close()


def test_one_substep_in_flight_per_stream(mechs):
    srv = H.Server(mechs)
    try:
        gri = mechs[1]
        y, P = _state(gri, 1, 2048, 2)
        sid = srv.open_stream(1, y, P)
        rid = srv.submit(sid, 1e-4, want_state=False)
        with pytest.raises(H.HpoError):
            srv.submit(sid, 1e-4, want_state=False)
        with pytest.raises(H.HpoError):
            srv.close_stream(sid)
        srv.wait(rid)
        srv.close_stream(sid)
    finally:
        srv.close()


def test_requests_arriving_together_share_launches(mechs):
    srv = H.Server(mechs, n_slots=1)
    try:
        h2 = mechs[0]
        sids = []
        for k in range(24):
            y, P = _state(h2, 0, 64, 40 + k)
            sids.append(srv.open_stream(0, y, P))
        rids = [srv.submit(s, 5e-5, want_state=False) for s in sids]
        infos = [srv.wait(r)[2] for r in rids]
        assert max(i["batch_requests"] for i in infos) > 1
        assert all(i["failed_cells"] == 0 for i in infos)
        for s in sids:
            srv.close_stream(s)
    finally:
        srv.close()


def test_short_poisson_load(mechs):
    srv = H.Server(mechs)
    try:
        r = SV.run_load(srv, mechs, rate_hz=20.0, duration_s=1.5, seed=3, substeps=2, max_cells=4096)
        assert r["substeps_completed"] == 2 * r["requests"]
        assert r["failed_cells"] == 0
        assert r["latency_ms_p99"] >= r["latency_ms_p50"] > 0
    finally:
        srv.close()


def test_sixteen_slots_under_load(mechs):
    """16 launch slots (16 HIP streams over the process's hardware queues) under a Poisson load
    with opens and closes from many client threads: every substep completes."""
    srv = H.Server(mechs, n_slots=16)
    try:
        r = SV.run_load(srv, mechs, rate_hz=40.0, duration_s=3.0, seed=5, substeps=2, max_cells=16384)
        print(r, flush=True)
        assert r["substeps_completed"] == 2 * r["requests"]
        assert r["failed_cells"] == 0
    finally:
        srv.close()


def test_device_order_changes_no_result(mechs):
    """Each launch's work-queue order is built on the device (hottest cells first, the default):
    every cell is integrated once, to the same result as in packing order, and as when the
    batch runs as consecutive launches of at most 1,000 cells (launch_cells)."""
    gri = mechs[1]
    y0, P = _state(gri, 1, 5000, 41)
    out = []
    for sort, lc in ((1, 0), (0, 0), (1, 1000)):
        srv = H.Server(mechs, n_slots=2, sort_cells=sort, launch_cells=lc)
        try:
            sid = srv.open_stream(1, y0, P)
            y, s, _ = srv.wait(srv.submit(sid, 6e-5, n=5000, n_state=gri.n_state))
            srv.close_stream(sid)
        finally:
            srv.close()
        assert (s == 0).all()
        out.append(y)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    assert not np.array_equal(out[0][:, 0], y0[:, 0])
