"""The wide-spread workload (T0 U[300,2500] K, phi log-U[0.3,3], dt log-U[1e-6,1e-3] s) and the
autotuner of the persistent kernel: per-cell results against the oracle, and results that do not
depend on the launch shape or the work-queue order the autotuner picks."""
import numpy as np
import pytest

import hpc_ode_amd as H
from hpc_ode_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gri(gri_path):
    import oracle as O
    return H.Mechanism(gri_path), O.Oracle(gri_path)


@pytest.mark.parametrize("lu_precision", [0, 1])
def test_uneven_cells_match_oracle(gri, lu_precision):
    """256 wide-spread cells at the stated bar (T and every mass fraction above 1e-10 within a
    relative 1e-6 of the oracle), judged against the oracle's round-off floor on the same cells
    (oracle/floor.py), plus a weighted sanity bound on every cell."""
    from oracle import floor as F
    m, o = gri
    y0, P, dt = WL.gri_uneven(m.species, m.W, 256, seed=31)
    fl = F.roundoff_floor(o, P, dt, y0, nthreads=8)
    yg, sg, stg = m.integrate(y0, P, dt, lu_precision=lu_precision, refine_sweeps=1)
    yo, so, sto = fl["y"], fl["status"], fl["stats"]
    assert (sg == 0).all() and (so == 0).all()
    print("floor", F.summary(fl), "gpu", F.compare(yg, sg, stg, fl), flush=True)
    F.check_against_floor(yg, sg, stg, fl)
    dT = np.abs(yg[:, 0] - yo[:, 0]) / yo[:, 0]
    assert dT.max() < 1e-5
    w = np.abs(yg - yo) / (1e-6 * np.abs(yo) + 1e-10)
    assert w.max() < 300.0
    assert abs(stg["nst"].mean() - sto["nst"].mean()) <= 0.03 * sto["nst"].mean()


def test_autotune_order_and_shape_do_not_change_results(gri):
    """Each cell is integrated by one wave with the same instruction sequence whatever the
    launch shape or queue order: results and counters are bit-identical."""
    m, _ = gri
    n = 4096
    y0, P, dt = WL.gri_uneven(m.species, m.W, n, seed=32)
    b = H.Batch(m, n)
    b.put(y0, P, dt)
    o = H.opts(lu_precision=1, refine_sweeps=1)
    rep = b.autotune(n, o, sample=1024)
    # 3 launch shapes x 3 queue orders of the wave-per-cell kernel, the block-per-cell kernel at
    # the best order, then (when a wave layout wins) the other FP32 Newton-matrix layout
    assert rep.n_candidates in (10, 11) and 0 <= rep.best < rep.n_candidates and rep.sample_cells == 1024
    cands = rep.candidates()
    assert [c["cell_layout"] for c in cands[:10]] == ["wave"] * 9 + ["block"]
    assert [c["lu_layout"] for c in cands[:10]] == ["inverse"] * 10
    if rep.n_candidates == 11:
        assert cands[10]["cell_layout"] == "wave" and cands[10]["lu_layout"] == "lu"
    assert o.waves_per_block in (1, 2, 4) and o.sort_cells == 2 and o.lu_layout in (0, 1)
    assert o.cell_layout in (1, 2)
    assert min(c["sample_ms"] for c in rep.candidates()) == rep.candidates()[rep.best]["sample_ms"]
    # tuning integrates a copy: the resident cells are unchanged
    y_before, _, _ = b.get(n)
    assert np.array_equal(y_before, y0)

    def run(opts):
        b.put(y0, P, dt)
        b.integrate(n, opts)
        return b.get(n)

    ya, sa, ca = run(o)
    yb, sb, cb = run(H.opts(lu_precision=1, refine_sweeps=1, waves_per_block=4, lu_layout=o.lu_layout,
                            cell_layout=o.cell_layout))
    yc, sc, cc = run(H.opts(lu_precision=1, refine_sweeps=1, waves_per_block=1, sort_cells=1,
                            lu_layout=o.lu_layout, cell_layout=o.cell_layout))
    assert (sa == 0).all()
    assert np.array_equal(ya, yb) and np.array_equal(ya, yc)
    for f in H.STATS_FIELDS:
        assert np.array_equal(ca[f], cb[f]) and np.array_equal(ca[f], cc[f]), f
    # after a temperature sort the tuned order is gone
    with pytest.raises(H.HpoError):
        b.integrate(n, o)
