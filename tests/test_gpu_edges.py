"""Edge cases of the launch: no cells at all, and ragged cell counts.

* ncells = 0 is a no-op on every entry point: nothing is written and the call succeeds.
* A cell's result does not depend on how many cells share its launch or where it sits: the
  same 300 cells integrated as one launch and as launches of 1, 3, 65 and 231 cells give the
  same states, statuses and step counts bit for bit, in both Newton-matrix precisions and in
  both cell layouts (one wave per cell on GRI-3.0, one workgroup per cell on nDodecane)."""
import ctypes as C
import os

import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu

MECH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hpc_ode_amd", "mechanisms")


def _batch_run(m, y0, P, dt, **kw):
    b = H.Batch(m, len(y0))
    b.put(y0, P, dt)
    b.integrate(len(y0), H.opts(**kw))
    return b.get(len(y0))


def test_zero_cells_is_a_no_op():
    m = H.Mechanism(os.path.join(MECH, "gri30.yaml"))
    y0, P, dt = WL.gri_plain(m.species, m.W, 4, seed=2)
    b = H.Batch(m, 4)
    b.put(y0, P, dt)
    for prec in (1, 0):
        t = b.integrate(0, H.opts(lu_precision=prec))
        assert t.ncells == 0 and t.fallback_cells == 0
    y, st, _ = b.get(4)
    assert np.array_equal(y, y0)
    # the device entry points with no cells
    L = H.lib()
    hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), 1 << 16) == 0
    try:
        o = H.opts(lu_precision=1)
        H._check(L.hpo_integrate_device(m.handle, 0, p, p, m.n_state, p, p, p, None, None, None, C.byref(o),
                                        None, None))
        H._check(L.hpo_integrate_device_noalloc(m.handle, 0, p, p, m.n_state, p, p, p, None, None, C.byref(o),
                                                None))
        assert hip.hipDeviceSynchronize() == 0
    finally:
        hip.hipFree(p)


@pytest.mark.parametrize("name,cell_layout", [("gri30.yaml", 1), ("nDodecane_Reitz.yaml", 2)])
@pytest.mark.parametrize("lu_precision", [1, 0])
def test_ragged_launches_match_one_launch(name, cell_layout, lu_precision):
    m = H.Mechanism(os.path.join(MECH, name))
    n = 300
    if name.startswith("gri"):
        y0, P, dt = WL.gri_plain(m.species, m.W, n, seed=12)
    else:
        y0, P, dt = WL.dodecane_air(m.species, m.W, n, seed=12)
    kw = dict(lu_precision=lu_precision, cell_layout=cell_layout)
    print(f"{name}: {n} cells, layout {cell_layout}, lu_precision {lu_precision}", flush=True)
    y, st, stats = _batch_run(m, y0, P, dt, **kw)
    assert (st == 0).all()
    cuts = [0, 1, 4, 69, n]
    parts = [_batch_run(m, y0[a:b], P[a:b], dt[a:b], **kw) for a, b in zip(cuts[:-1], cuts[1:])]
    assert np.array_equal(np.concatenate([p[0] for p in parts]), y)
    assert np.array_equal(np.concatenate([p[1] for p in parts]), st)
    assert np.array_equal(np.concatenate([p[2]["nst"] for p in parts]), stats["nst"])
