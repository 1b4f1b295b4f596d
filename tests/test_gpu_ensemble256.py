"""The 256-member ensemble of 1-D RF discharges (grid sizes 128-1024 cells,
10^4-5*10^5 particles per species, 100-400 V, 0.1-2 Torr, per-instance
seeds: ensemble.draw_instances(256, seed=0), the draw the benchmark uses)
advanced as one batch for one captured sort cycle (32 steps), against

* every member stepped alone with iskra_step under a different tile and chunk
  shape than in the batch -- particle states compared by id, rho, phi and
  the collision counts bit for bit (ionisation products are numbered by their
  parents' ids, so no result depends on the work decomposition);
* three members against the CPU twin (oracle.MirSim) from the same initial
  particles, bit for bit by id."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(32, 1024), (64, 2048), (64, 4096), (128, 4096)]


def _by_id(st):
    live = st["cell"] >= 0
    o = np.argsort(st["id"][live], kind="stable")
    return {k: v[live][o] for k, v in st.items()}


def _same(a, b, what):
    for s in (0, 1):
        sa, sb = a[s], b[s]
        assert len(sa["id"]) == len(sb["id"]), (what, s, len(sa["id"]), len(sb["id"]))
        for k in ("cell", "ox", "vx", "vy", "vz", "id"):
            assert np.array_equal(sa[k].view(np.uint32), sb[k].view(np.uint32)), (what, s, k)


def _twin(orc, inst, ctx, tile):
    """MirSim of one ensemble member from the member's device state."""
    from iskra_amd import problems as P
    cells = inst["cells"]
    p = P.ccp_1d(cells=cells, electrons=inst["particles"], ions=inst["particles"],
                 pressure_torr=inst["torr"], volts=inst["volts"])
    nodes = cells + 1
    sim = orc.MirSim(nodes, p["dx"], T=tile, sort_every=16, seed=inst["seed"])
    for sid, (q, m) in enumerate([(-p["q"], p["me"]), (p["q"], p["mi"])]):
        st = ctx.get_state(sid)
        sp = orc.MirSpecies(1, int(2.0 * inst["particles"]) + 1024, q, m, p["w0"], bnd=(2, 2))
        sp.set_state(st["cell"], st["ox"], None, st["vx"], st["vy"], st["vz"], st["id"])
        sim.add_species(sp)
    tn = np.full(nodes, p["ng"])
    xs = p["xs"]
    sim.add_mcc(0, 1, [0, 2, 3], xs["energies"], xs["electron"], p["mi"], 0.0, 300.0, tn)
    sim.add_mcc(1, None, [1], [0.0], xs["ion"], p["mi"], 0.0, 300.0, tn)
    tags = np.zeros(nodes, np.int8)
    tags[0], tags[-1] = 1, 2
    sim.set_poisson(p["eps0"], tags, 0)
    sim.wf = {1: (0.0, inst["volts"], 13.56e6, 0.0)}
    return sim


def test_ensemble_256_members_equal_single_runs(gpu_ready, orc):
    from iskra_amd import ensemble as E, _lib as L
    from iskra_amd._lib import Batch
    insts = E.draw_instances(256, seed=0)
    steps = 32
    shapes_b = [SHAPES[k % 4] for k in range(256)]
    shapes_s = [SHAPES[(k + 1) % 4] for k in range(256)]
    order = np.argsort([i["particles"] for i in insts])
    twin_k = [int(order[0]), int(order[1]), int(order[2])]
    twins = {}
    for k in twin_k:
        c, _ = E.ccp_context(insts[k], tile=shapes_b[k][0], chunk=shapes_b[k][1], sort_every=16)
        twins[k] = _twin(orc, insts[k], c, shapes_b[k][0])
        c.close()
    ctxs, dt = E.build(insts, shapes_b, sort_every=16)
    b = Batch(ctxs)
    try:
        total = sum(c.count(0) + c.count(1) for c in ctxs)
        print(f"256 members, {total} particles", flush=True)
        assert total > 3e7
        b.step(dt, steps)
        assert b.graphs_launched() == 1
        ions = 0
        for k, inst in enumerate(insts):
            got = [_by_id(ctxs[k].get_state(s)) for s in (0, 1)]
            c, _ = E.ccp_context(inst, tile=shapes_s[k][0], chunk=shapes_s[k][1],
                                 sort_every=16)
            try:
                c.step(dt, steps)
                want = [_by_id(c.get_state(s)) for s in (0, 1)]
                _same(got, want, f"member {k}")
                assert np.array_equal(ctxs[k].field(L.F_RHO).view(np.uint64),
                                      c.field(L.F_RHO).view(np.uint64)), k
                assert np.array_equal(ctxs[k].field(L.F_PHI).view(np.uint64),
                                      c.field(L.F_PHI).view(np.uint64)), k
                nu = ctxs[k].mcc_counts(0, 3)
                assert np.array_equal(nu, c.mcc_counts(0, 3)), k
                ions += int(nu[2].sum())
            finally:
                c.close()
            if k in twins:
                sim = twins[k]
                sim.step(dt, steps)
                tw = []
                for sp in sim.species:
                    n = sp.np
                    st = {f: getattr(sp, f)[:n].copy() for f in ("cell", "ox", "vx", "vy", "vz",
                                                                "id")}
                    tw.append(_by_id(st))
                _same(got, tw, f"member {k} vs CPU twin")
                assert np.array_equal(ctxs[k].field(L.F_PHI), sim.phi), k
            if k % 64 == 63:
                print(f"members 0..{k} equal", flush=True)
        assert ions > 0
    finally:
        E.close_all(ctxs, b)
