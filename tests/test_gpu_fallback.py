"""The FP64 fallback of the FP32 Newton matrix (hpo_opts::fp64_fallback, on by default): a cell
whose FP32 run ends in a solver failure (convergence or error-test failure, too much work, too
much accuracy) keeps its input row, and a second launch on the same stream re-integrates it in
FP64 from there. So with the fallback each cell's result is either its FP32 result (when that
succeeded) or, bit for bit, the plain FP64 kernel's result for the cell.

A step limit (maxsteps) below some cells' step counts makes those cells fail with
TOO_MUCH_WORK in FP32 and again in FP64 (where the FP64 result is the state after the same
number of FP64 steps), which exercises the mechanism on chosen cells without forcing a
numerical failure."""
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
