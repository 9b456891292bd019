"""hpo_integrate_device captured into a HIP graph: every replay, with eager integrations and
host synchronisations between the replays, equals the eager result bit for bit.

The launch enqueues a counter-reset kernel and the BDF kernel (no memset node). The HIP calls
go through the runtime libhpc_ode.so links, so device pointers and the stream are that
runtime's own."""
import ctypes as C

import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
CAPTURE_RELAXED = 2


def _hip():
    H.lib()   # loads libamdhip64.so.7 as a dependency of libhpc_ode.so
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    vp = C.c_void_p
    hip.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
    hip.hipStreamCreate.argtypes = [C.POINTER(vp)]
    hip.hipStreamDestroy.argtypes = [vp]
    hip.hipStreamSynchronize.argtypes = [vp]
    hip.hipStreamBeginCapture.argtypes = [vp, C.c_int]
    hip.hipStreamEndCapture.argtypes = [vp, C.POINTER(vp)]
    hip.hipGraphInstantiate.argtypes = [C.POINTER(vp), vp, vp, C.c_char_p, C.c_size_t]
    hip.hipGraphLaunch.argtypes = [vp, vp]
    hip.hipGraphExecDestroy.argtypes = [vp]
    hip.hipGraphDestroy.argtypes = [vp]
    return hip


def _ok(rc, what):
    assert rc == 0, f"{what} returned hipError {rc}"


@pytest.mark.parametrize("lu_precision", [1, 0])
def test_graph_replays_match_eager(gri_path, lu_precision):
    hip = _hip()
    m = H.Mechanism(gri_path)
    n, ns = 512, m.n_state
    y0, P, dt = WL.gri_plain(m.species, m.W, n, seed=31)
    y0 = np.ascontiguousarray(y0)
    bufs = {}

    def dev(name, nbytes):
        p = C.c_void_p()
        _ok(hip.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        bufs[name] = p
        return p

    try:
        d_y0, d_ye, d_yg = dev("y0", y0.nbytes), dev("ye", y0.nbytes), dev("yg", y0.nbytes)
        d_P, d_dt = dev("P", P.nbytes), dev("dt", dt.nbytes)
        d_se, d_sg = dev("se", 4 * n), dev("sg", 4 * n)
        _ok(hip.hipMemcpy(d_y0, y0.ctypes.data, y0.nbytes, H2D), "copy y0")
        _ok(hip.hipMemcpy(d_P, P.ctypes.data, P.nbytes, H2D), "copy P")
        _ok(hip.hipMemcpy(d_dt, dt.ctypes.data, dt.nbytes, H2D), "copy dt")
        st = C.c_void_p()
        _ok(hip.hipStreamCreate(C.byref(st)), "hipStreamCreate")
        bufs["stream"] = st
        o = H.opts(lu_precision=lu_precision)
        L = H.lib()

        def eager():
            H._check(L.hpo_integrate_device(m.handle, n, d_y0, d_ye, ns, d_P, d_dt, d_se, None, None, None,
                                            C.byref(o), st, None))
            _ok(hip.hipStreamSynchronize(st), "sync")
            y = np.empty_like(y0)
            s = np.empty(n, np.int32)
            _ok(hip.hipMemcpy(y.ctypes.data, d_ye, y.nbytes, D2H), "copy y")
            _ok(hip.hipMemcpy(s.ctypes.data, d_se, s.nbytes, D2H), "copy status")
            return y, s

        y_ref, s_ref = eager()            # also sizes the mechanism's scratch before capture
        assert (s_ref == 0).all()

        graph, exe = C.c_void_p(), C.c_void_p()
        _ok(hip.hipStreamBeginCapture(st, CAPTURE_RELAXED), "begin capture")
        rc = L.hpo_integrate_device(m.handle, n, d_y0, d_yg, ns, d_P, d_dt, d_sg, None, None, None,
                                    C.byref(o), st, None)
        _ok(hip.hipStreamEndCapture(st, C.byref(graph)), "end capture")
        H._check(rc)
        _ok(hip.hipGraphInstantiate(C.byref(exe), graph, None, None, 0), "instantiate")
        try:
            for rep in range(3):
                _ok(hip.hipMemset(d_yg, 0, y0.nbytes), "clear output")
                _ok(hip.hipMemset(d_sg, 0xFF, 4 * n), "clear status")
                _ok(hip.hipGraphLaunch(exe, st), "graph launch")
                _ok(hip.hipStreamSynchronize(st), "sync")
                y = np.empty_like(y0)
                s = np.empty(n, np.int32)
                _ok(hip.hipMemcpy(y.ctypes.data, d_yg, y.nbytes, D2H), "copy y")
                _ok(hip.hipMemcpy(s.ctypes.data, d_sg, s.nbytes, D2H), "copy status")
                assert np.array_equal(y, y_ref), f"replay {rep}"
                assert np.array_equal(s, s_ref), f"replay {rep}"
                ye, se = eager()          # eager work between replays
                assert np.array_equal(ye, y_ref) and np.array_equal(se, s_ref)
        finally:
            hip.hipGraphExecDestroy(exe)
            hip.hipGraphDestroy(graph)
    finally:
        st = bufs.pop("stream", None)
        if st is not None:
            hip.hipStreamDestroy(st)
        for p in bufs.values():
            hip.hipFree(p)
