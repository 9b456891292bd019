"""The FP64 fallback of the FP32 Newton matrix (hpo_opts::fp64_fallback, on by default): a cell
whose FP32 run ends in a solver failure (convergence or error-test failure, too much work, too
much accuracy) is re-integrated in FP64 from its input row by a second launch on the same
stream (an in-place call first copies its input rows). So with the fallback each cell's result
is either its FP32 result (when that succeeded) or, bit for bit, the plain FP64 kernel's result
for the cell, whichever entry point ran it: the batch (in place), device arrays (separate
input and output), the serving loop (in place in its slots). The no-allocation entry point runs
without the fallback in place and with it otherwise.

A step limit (maxsteps) below some cells' step counts makes those cells fail with
TOO_MUCH_WORK in FP32 and again in FP64 (where the FP64 result is the state after the same
number of FP64 steps), which exercises the mechanism on chosen cells without forcing a
numerical failure."""
import ctypes as C
import os

import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu

MECH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms")
TOO_MUCH_WORK = -1


def _run(m, y0, P, dt, **kw):
    b = H.Batch(m, len(y0))
    b.put(y0, P, dt)
    t = b.integrate(len(y0), H.opts(**kw))
    y, st, stats = b.get(len(y0))
    return y, st, stats, t


@pytest.mark.parametrize("name,n,cell_layout", [("gri30.yaml", 512, 1), ("nDodecane_Reitz.yaml", 64, 2)])
def test_failed_fp32_cells_get_the_fp64_result(name, n, cell_layout):
    m = H.Mechanism(os.path.join(MECH, name))
    if name.startswith("gri"):
        y0, P, dt = WL.gri_plain(m.species, m.W, n, seed=9)
    else:
        y0, P, dt = WL.dodecane_air(m.species, m.W, n, seed=9)
    y32, s32, st32, _ = _run(m, y0, P, dt, lu_precision=1, cell_layout=cell_layout, fp64_fallback=0)
    assert (s32 == 0).all()
    limit = int(np.median(st32["nst"]))   # about half the cells need more steps than this
    print(f"{name}: {n} cells, maxsteps {limit}", flush=True)
    yf, sf, stf, tf = _run(m, y0, P, dt, lu_precision=1, cell_layout=cell_layout, maxsteps=limit)
    y64, s64, st64, _ = _run(m, y0, P, dt, lu_precision=0, cell_layout=cell_layout, maxsteps=limit)
    yn, sn, stn, _ = _run(m, y0, P, dt, lu_precision=1, cell_layout=cell_layout, maxsteps=limit, fp64_fallback=0)
    failed = sn == TOO_MUCH_WORK
    assert 0 < failed.sum() < n and ((sn == 0) | failed).all()
    assert tf.fallback_cells == failed.sum()
    # cells the FP32 run finished: their FP32 result, as without the fallback
    assert np.array_equal(yf[~failed], yn[~failed]) and (sf[~failed] == 0).all()
    assert np.array_equal(stf["nst"][~failed], stn["nst"][~failed])
    # the others: the FP64 kernel's result, state, status and counters alike
    assert np.array_equal(yf[failed], y64[failed])
    assert np.array_equal(sf[failed], s64[failed])
    for f in H.STATS_FIELDS:
        assert np.array_equal(stf[f][failed], st64[f][failed]), f


def test_no_failure_no_change():
    """Without failures the fallback launch returns at once and the results are the FP32 ones."""
    m = H.Mechanism(os.path.join(MECH, "gri30.yaml"))
    y0, P, dt = WL.gri_plain(m.species, m.W, 1024, seed=10)
    ya, sa, sta, ta = _run(m, y0, P, dt, lu_precision=1)
    yb, sb, stb, tb = _run(m, y0, P, dt, lu_precision=1, fp64_fallback=0)
    assert (sa == 0).all() and ta.fallback_cells == 0
    assert np.array_equal(ya, yb) and np.array_equal(sta["nst"], stb["nst"])


def _hip():
    H.lib()   # loads libamdhip64.so.7 as a dependency of libhpc_ode.so
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    vp = C.c_void_p
    hip.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def test_fallback_through_device_arrays_and_serving():
    """The same limited-step GRI cells through hpo_integrate_device (separate input and output
    arrays, then in place), hpo_integrate_device_noalloc (separate: with the fallback; in place:
    without it) and the serving loop, against the batch's results."""
    m = H.Mechanism(os.path.join(MECH, "gri30.yaml"))
    n, ns = 384, m.n_state
    y0, P, dt = WL.gri_plain(m.species, m.W, n, seed=21)
    y0 = np.ascontiguousarray(y0)
    _, _, st32, _ = _run(m, y0, P, dt, lu_precision=1, fp64_fallback=0)
    limit = int(np.median(st32["nst"]))
    yf, sf, _, tf = _run(m, y0, P, dt, lu_precision=1, maxsteps=limit)
    yn, sn, _, _ = _run(m, y0, P, dt, lu_precision=1, maxsteps=limit, fp64_fallback=0)
    failed = sn == TOO_MUCH_WORK
    assert 0 < failed.sum() < n and tf.fallback_cells == failed.sum()
    assert not np.array_equal(yf[failed], yn[failed])

    hip, L = _hip(), H.lib()
    bufs = []

    def dev(nbytes):
        p = C.c_void_p()
        assert hip.hipMalloc(C.byref(p), nbytes) == 0
        bufs.append(p)
        return p

    def put(p, a):
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0

    def get(p, shape, dtype):
        a = np.empty(shape, dtype)
        assert hip.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0
        return a

    o = H.opts(lu_precision=1, maxsteps=limit)
    try:
        d_y0, d_y, d_P, d_dt, d_s = dev(y0.nbytes), dev(y0.nbytes), dev(P.nbytes), dev(dt.nbytes), dev(4 * n)
        put(d_y0, y0), put(d_P, P), put(d_dt, dt)
        print("hpo_integrate_device, separate arrays", flush=True)
        H._check(L.hpo_integrate_device(m.handle, n, d_y0, d_y, ns, d_P, d_dt, d_s, None, None, None,
                                        C.byref(o), None, None))
        assert hip.hipDeviceSynchronize() == 0
        assert np.array_equal(get(d_y, (n, ns), np.float64), yf) and np.array_equal(get(d_s, n, np.int32), sf)
        print("hpo_integrate_device, in place", flush=True)
        put(d_y, y0)
        H._check(L.hpo_integrate_device(m.handle, n, d_y, d_y, ns, d_P, d_dt, d_s, None, None, None,
                                        C.byref(o), None, None))
        assert hip.hipDeviceSynchronize() == 0
        assert np.array_equal(get(d_y, (n, ns), np.float64), yf) and np.array_equal(get(d_s, n, np.int32), sf)
        print("hpo_integrate_device_noalloc, separate arrays", flush=True)
        H._check(L.hpo_integrate_device_noalloc(m.handle, n, d_y0, d_y, ns, d_P, d_dt, d_s, None, None,
                                                C.byref(o), None))
        assert hip.hipDeviceSynchronize() == 0
        assert np.array_equal(get(d_y, (n, ns), np.float64), yf) and np.array_equal(get(d_s, n, np.int32), sf)
        print("hpo_integrate_device_noalloc, in place", flush=True)
        put(d_y, y0)
        H._check(L.hpo_integrate_device_noalloc(m.handle, n, d_y, d_y, ns, d_P, d_dt, d_s, None, None,
                                                C.byref(o), None))
        assert hip.hipDeviceSynchronize() == 0
        assert np.array_equal(get(d_y, (n, ns), np.float64), yn) and np.array_equal(get(d_s, n, np.int32), sn)
    finally:
        hip.hipDeviceSynchronize()
        for p in bufs:
            hip.hipFree(p)

    print("serving loop", flush=True)
    srv = H.Server([m], n_slots=2, lu_precision=1, maxsteps=limit)
    try:
        sid = srv.open_stream(0, y0, P)
        ys, ss, info = srv.wait(srv.submit(sid, float(dt[0]), n=n, n_state=ns))
        srv.close_stream(sid)
    finally:
        srv.close()
    assert (dt == dt[0]).all()
    assert info["failed_cells"] == (sf != 0).sum()
    assert np.array_equal(ys, yf) and np.array_equal(ss, sf)
