"""hpo_integrate_device_noalloc (the entry point that allocates nothing) against
hpo_integrate_device on the same device arrays: bit-identical states, statuses and work
counters (one wave per cell runs the same instructions whichever scratch it uses), on two calls
in a row, the second with fewer cells, in both Newton-matrix precisions; and the refusal when
the mechanism's tables are not on the device yet."""
import ctypes as C

import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    H.lib()   # loads libamdhip64.so.7 as a dependency of libhpc_ode.so
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    vp = C.c_void_p
    hip.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(vp)]
    hip.hipStreamDestroy.argtypes = [vp]
    hip.hipStreamSynchronize.argtypes = [vp]
    return hip


def _ok(rc, what):
    assert rc == 0, f"{what} returned hipError {rc}"


@pytest.mark.parametrize("lu_precision", [1, 0])
def test_noalloc_matches_allocating_entry(gri_path, lu_precision):
    hip = _hip()
    L = H.lib()
    m = H.Mechanism(gri_path)
    ns = m.n_state
    n = 2048
    y0, P, dt = WL.gri_plain(m.species, m.W, n, seed=77)
    y0 = np.ascontiguousarray(y0)
    bufs = []

    def dev(nbytes):
        p = C.c_void_p()
        _ok(hip.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        bufs.append(p)
        return p

    def get(p, shape, dtype):
        a = np.empty(shape, dtype)
        _ok(hip.hipMemcpy(a.ctypes.data, p, a.nbytes, D2H), "copy back")
        return a

    st = C.c_void_p()
    _ok(hip.hipStreamCreate(C.byref(st)), "hipStreamCreate")
    try:
        d_y0, d_ya, d_yb = dev(y0.nbytes), dev(y0.nbytes), dev(y0.nbytes)
        d_P, d_dt = dev(P.nbytes), dev(dt.nbytes)
        d_sa, d_sb = dev(4 * n), dev(4 * n)
        d_ca, d_cb = dev(4 * n * H.NSTATS), dev(4 * n * H.NSTATS)
        _ok(hip.hipMemcpy(d_y0, y0.ctypes.data, y0.nbytes, H2D), "copy y0")
        _ok(hip.hipMemcpy(d_P, P.ctypes.data, P.nbytes, H2D), "copy P")
        _ok(hip.hipMemcpy(d_dt, dt.ctypes.data, dt.nbytes, H2D), "copy dt")
        o = H.opts(lu_precision=lu_precision, refine_sweeps=1)
        # the tables are not on the device before any allocating call
        rc = L.hpo_integrate_device_noalloc(m.handle, n, d_y0, d_yb, ns, d_P, d_dt, d_sb, d_cb, None,
                                            C.byref(o), st)
        assert rc == H.E_INVALID and b"tables are not on the current device" in L.hpo_last_error()
        for k, count in enumerate((n, 700)):
            print(f"call {k}: {count} cells, lu_precision {lu_precision}", flush=True)
            H._check(L.hpo_integrate_device(m.handle, count, d_y0, d_ya, ns, d_P, d_dt, d_sa, d_ca, None, None,
                                            C.byref(o), st, None))
            H._check(L.hpo_integrate_device_noalloc(m.handle, count, d_y0, d_yb, ns, d_P, d_dt, d_sb, d_cb, None,
                                                    C.byref(o), st))
            _ok(hip.hipStreamSynchronize(st), "sync")
            ya, yb = get(d_ya, (n, ns), np.float64)[:count], get(d_yb, (n, ns), np.float64)[:count]
            sa, sb = get(d_sa, n, np.int32)[:count], get(d_sb, n, np.int32)[:count]
            ca, cb = (get(d_ca, (n, H.NSTATS), np.int32)[:count], get(d_cb, (n, H.NSTATS), np.int32)[:count])
            assert (sa == 0).all()
            assert np.array_equal(ya, yb) and np.array_equal(sa, sb) and np.array_equal(ca, cb)
    finally:
        _ok(hip.hipStreamSynchronize(st), "sync")
        for p in bufs:
            hip.hipFree(p)
        hip.hipStreamDestroy(st)
