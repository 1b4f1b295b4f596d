"""DSMC on the device (Chemistry/src/dsmc.jl:81-142 with the elastic
collision of dsmc.jl:32-79) against the CPU twin bit for bit: one call, and
the time loop with DSMC inside iskra_step; the configurations the reference
cannot run are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ME, MO = 9.10938e-31, 16 * 1.6726e-27


def _setup(orc, nodes=33, dx=1e-3, per_cell=8, w0=2e16, vth=(1e6, 1e3), q=(-1.602e-19, 0.0),
           masses=(ME, MO), seed=7):
    from iskra_amd import Context, _lib as L
    ctx = Context()
    ctx.grid(2, (nodes, nodes), (dx, dx))
    ctx.tuning(tile=8, chunk=1024, margin=2, sort_every=4)
    ctx.rng(seed)
    sim = orc.MirSim(nodes, dx, T=8, sort_every=4, ny=nodes, dy=dx, seed=seed)
    n = per_cell * (nodes - 1) ** 2
    L_ = (nodes - 1) * dx
    for k in range(2):
        s = ctx.species(q[k], masses[k], w0, n, track_ids=True)
        ctx.boundary(s, [L.BND_DISCARD, L.BND_DISCARD])
        ctx.sample_maxwellian(s, n, [L_, L_], [vth[k]] * 3, seed=seed, stream=k)
        st = ctx.get_state(s)
        sp = orc.MirSpecies(2, n, q[k], masses[k], w0, bnd=(2, 2))
        sp.set_state(st["cell"], st["ox"], st["oy"], st["vx"], st["vy"], st["vz"], st["id"])
        sim.add_species(sp)
    return ctx, sim, n


def _assert_species_equal(ctx, sim, sid, what):
    st = ctx.get_state(sid)
    sp = sim.species[sid]
    assert len(st["cell"]) == sp.np, (what, len(st["cell"]), sp.np)
    for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id"):
        a, b = st[k], getattr(sp, k)[:sp.np]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (what, sid, k)


TAB_G = [1e5, 1e6, 3e6]
TAB_S = [1e-19, 2e-19, 1e-19]


def test_dsmc_call_bitexact(gpu_ready, orc):
    """Three calls of iskra_dsmc_run (electrons on oxygen atoms, ~10 trials
    per cell) against mir_dsmc: particle order (exact-cell sort), velocities,
    event counts, per-node counts and the carried remainders."""
    ctx, sim, n = _setup(orc)
    try:
        did = ctx.dsmc(0, 1, TAB_G, TAB_S)
        k = sim.add_dsmc(0, 1, TAB_G, TAB_S)
        sgmax = 2e-19 * 1e6
        na = 8 * 2e16 / (1e-3 * 1e-3)
        dt = 10.0 / (na * 8 * sgmax)  # Nc ~ 10 in a cell holding 8 + 8
        for call in range(3):
            ev = ctx.dsmc_run(did, dt)
            ev_m = sim.dsmc_run(k, dt)
            assert ev == ev_m, (call, ev, ev_m)
            assert ev > 1000
            for s in range(2):
                _assert_species_equal(ctx, sim, s, f"call {call}")
        assert np.array_equal(ctx.dsmc_counts(did), sim.dsmcs[k]["nu"])
        assert np.array_equal(ctx.dsmc_counts(did, reset=True), sim.dsmcs[k]["nu"])
        assert not ctx.dsmc_counts(did).any()
    finally:
        ctx.close()


def test_dsmc_in_time_loop_bitexact(gpu_ready, orc):
    """Two neutral gases (no field) colliding by DSMC inside iskra_step for 12
    steps, streaming and leaving through discard walls: every 4 steps the
    particle states equal the twin's bit for bit (its DSMC runs before the
    push, as iskra_step orders them)."""
    from iskra_amd import _lib as L
    ctx, sim, n = _setup(orc, vth=(3e3, 1e3), q=(0.0, 0.0), masses=(MO, 2 * MO), w0=4e17)
    nodes, dx = 33, 1e-3
    tags = np.zeros((nodes, nodes), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    try:
        ctx.poisson(L.POISSON_PCG, 8.854e-12, tags.ravel(), rtol=1e-10)
        ctx.tag_value(1, 0.0)
        sim.set_poisson_2d(8.854e-12, tags, [0.0, 0.0])
        did = ctx.dsmc(0, 1, [1e2, 1e3, 1e4], [3e-19, 4e-19, 2e-19])
        k = sim.add_dsmc(0, 1, [1e2, 1e3, 1e4], [3e-19, 4e-19, 2e-19])
        dt = 2e-8
        total = 0
        prev = sim.dsmcs[k]["nu"].copy()
        for it in range(12):
            ctx.step(dt, 1)
            sim.advance(dt)
            assert np.array_equal(ctx.field(L.F_RHO), sim.rho), it
            # the step's own counts (the reference's nuDSMC, dsmc.jl:124-141)
            assert np.array_equal(ctx.dsmc_counts_last(did), sim.dsmcs[k]["nu"] - prev), it
            prev = sim.dsmcs[k]["nu"].copy()
            sim.set_phi(ctx.field(L.F_PHI))
            sim.end_step(dt)
            total += sim.dsmcs[k]["last"]
            if it % 4 == 3:
                for s in range(2):
                    _assert_species_equal(ctx, sim, s, f"step {it + 1}")
        assert total > 5000
        assert np.array_equal(ctx.dsmc_counts(did), sim.dsmcs[k]["nu"])
        assert sim.species[0].np < n  # particles left through the walls
    finally:
        ctx.close()


def test_dsmc_refusals(gpu_ready):
    from iskra_amd import Context, _lib as L
    from iskra_amd import chemistry as ch, particle_in_cell as pic, regular_grids as rg
    ctx = Context()
    try:
        ctx.grid(2, (9, 9), (1e-3, 1e-3))
        a = ctx.species(-1.0, ME, 1.0, 64)
        b = ctx.species(0.0, MO, 2.0, 64)
        with pytest.raises(L.IskraError) as ei:
            ctx.dsmc(a, b, TAB_G, TAB_S)  # unequal weights
        assert ei.value.code == L.E_UNSUPPORTED
        with pytest.raises(L.IskraError) as ei:
            ctx.dsmc(a, a, [2.0, 1.0], [1e-19, 1e-19])  # speeds must increase
        assert ei.value.code == L.E_ARG
        ctx.grid(1, (9,), (1e-3,))
        c = ctx.species(-1.0, ME, 1.0, 64)
        with pytest.raises(L.IskraError) as ei:
            ctx.dsmc(c, c, TAB_G, TAB_S)
        assert ei.value.code == L.E_UNSUPPORTED
    finally:
        ctx.close()
    # host layer: two reactions in one interaction, and ionisation, are refused
    g = rg.create_uniform_grid(np.arange(9) * 1e-3, np.arange(9) * 1e-3)
    e = pic.create_kinetic_species("e-", 64, -1.602e-19, ME, 1.0)
    o = pic.create_kinetic_species("O", 64, 0.0, MO, 1.0)
    sig = ch.CrossSection(TAB_G, TAB_S)
    for rl in (ch.reactions([(sig, [e, o], [o, e]), (sig, [e, o], [o, e])]),
               ch.reactions([(sig, [e, o], [o, e, e])])):
        inter = ch.dsmc(rl)
        cfg = pic.Config(grid=g, species=[e, o], interactions=[inter],
                         solver=None, pusher=pic.create_boris_pusher())
        with pytest.raises(L.IskraError) as ei:
            pic.bind(cfg)
        assert ei.value.code == L.E_UNSUPPORTED
        if cfg.ctx is not None:
            cfg.ctx.close()
