"""Batched ensembles (iskra_batch_*): every member of a batch of uneven 1-D
RF discharges (different grids, particle counts, voltages, pressures, seeds
and tile shapes) must end bit for bit where the same context ends when it is
stepped alone with iskra_step, with the batch's steps captured in hipGraphs or
launched one by one, in one call or several."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TUNINGS = [(32, 1024), (64, 2048), (16, 1024), (64, 1024), (128, 2048), (32, 2048)]


def _insts():
    from iskra_amd import ensemble as E
    return E.draw_instances(6, seed=3, cells=(64, 256), particles=(2000, 8000))


def _compare(ca, cb, what):
    from iskra_amd import _lib as L
    for s in (0, 1):
        sa, sb = ca.get_state(s), cb.get_state(s)
        assert len(sa["cell"]) == len(sb["cell"]), (what, s)
        for k in ("cell", "ox", "vx", "vy", "vz", "id"):
            assert np.array_equal(sa[k].view(np.uint32), sb[k].view(np.uint32)), (what, s, k)
    assert np.array_equal(ca.field(L.F_PHI), cb.field(L.F_PHI)), what
    assert np.array_equal(ca.field(L.F_RHO), cb.field(L.F_RHO)), what
    for m, nc in ((0, 3), (1, 1)):
        assert np.array_equal(ca.mcc_counts(m, nc), cb.mcc_counts(m, nc)), (what, m)


def test_batch_members_equal_single_runs(gpu_ready):
    from iskra_amd import ensemble as E
    from iskra_amd._lib import Batch
    insts = _insts()
    steps = 40
    ctxs, dt = E.build(insts, TUNINGS, sort_every=8)
    b = Batch(ctxs)
    b.step(dt, steps)
    assert b.graphs_launched() == 2  # two 16-step cycles captured/replayed, 8 steps eager
    singles = []
    for k, inst in enumerate(insts):
        c, _ = E.ccp_context(inst, tile=TUNINGS[k][0], chunk=TUNINGS[k][1], sort_every=8)
        c.step(dt, steps)
        singles.append(c)
    ions = 0
    for k in range(len(insts)):
        _compare(ctxs[k], singles[k], f"graph instance {k}")
        ions += int(ctxs[k].mcc_counts(0, 3)[2].sum())
        assert ctxs[k].stats()["steps"] == steps
    assert ions > 0  # ionisation products were appended
    # the same ensemble without graphs, in two calls of uneven length
    ctxs2, _ = E.build(insts, TUNINGS, sort_every=8)
    b2 = Batch(ctxs2)
    b2.graph(False)
    b2.step(dt, 17)
    b2.step(dt, steps - 17)
    assert b2.graphs_launched() == 0
    for k in range(len(insts)):
        _compare(ctxs2[k], singles[k], f"eager instance {k}")
    E.close_all(ctxs, b)
    E.close_all(ctxs2, b2)
    E.close_all(singles)


def test_one_member_batch_is_refused_under_a_profiler_tool(gpu_ready, monkeypatch):
    """iskra_batch_step refuses a one-member batch while a rocprofiler-sdk
    tool is loaded (ROCP_TOOL_LIBRARIES set) before it queues any work; the
    same batch steps normally without it."""
    from iskra_amd import _lib as L
    from iskra_amd import ensemble as E
    from iskra_amd._lib import Batch
    ctxs, dt = E.build(_insts()[:1], TUNINGS[:1], sort_every=8)
    b = Batch(ctxs)
    monkeypatch.setenv("ROCP_TOOL_LIBRARIES", "set-by-the-test")
    with pytest.raises(L.IskraError) as e:
        b.step(dt, 4)
    assert e.value.code == L.E_UNSUPPORTED
    assert ctxs[0].stats()["steps"] == 0
    monkeypatch.delenv("ROCP_TOOL_LIBRARIES")
    b.step(dt, 4)
    assert ctxs[0].stats()["steps"] == 4
