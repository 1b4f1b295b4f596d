"""GPU parity of whole steps: MCC and the 1-D time loop against the oracle's
CPU twin (oracle.MirSim), bit for bit, given the same Philox draws."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ccp_setup(nodes=129, ne=20000, ni=20000, seed=77, Te_eV=12.0, T=64):
    from iskra_amd import Context, _lib as L, problems as P
    import oracle as O
    p = P.ccp_1d(cells=nodes - 1, electrons=ne, ions=ni)
    rng = np.random.default_rng(seed)
    ctx = Context()
    ctx.grid(1, (nodes,), (p["dx"],))
    ctx.tuning(tile=T, sort_every=4)
    ctx.rng(seed)
    cap = 3 * ne
    e = ctx.species(-p["q"], p["me"], p["w0"], cap, track_ids=True)
    i = ctx.species(p["q"], p["mi"], p["w0"], cap, track_ids=True)
    ctx.boundary(e, [L.BND_DISCARD])
    ctx.boundary(i, [L.BND_DISCARD])
    ve = P.thermal_speed(Te_eV * P.QE / P.KB, P.ME)
    xe = rng.random(ne) * p["L"]
    xi = rng.random(ni) * p["L"]
    ctx.upload(e, xe[:, None], rng.standard_normal((ne, 3)) * ve)
    ctx.upload(i, xi[:, None], rng.standard_normal((ni, 3)) * p["vi"])
    xs = p["xs"]
    me_id = ctx.mcc(e, [L.COLL_ELASTIC_ISOTROPIC, L.COLL_EXCITATION, L.COLL_IONIZATION],
                    xs["energies"], xs["electron"], ion=i, target_m=p["mi"], target_q=0.0,
                    target_T=p["Tg"])
    mi_id = ctx.mcc(i, [L.COLL_ELASTIC_BACKWARD], [0.0], xs["ion"], ion=-1, target_m=p["mi"],
                    target_q=0.0, target_T=p["Tg"])
    tn = np.full(nodes, p["ng"])
    ctx.set_field(L.F_TARGET, tn, index=me_id)
    ctx.set_field(L.F_TARGET, tn, index=mi_id)
    tags = np.zeros(nodes, np.int8)
    tags[0], tags[-1] = 1, 2
    ctx.poisson(L.POISSON_DIRECT, p["eps0"], tags)
    ctx.tag_waveform(1, 0.0, p["volts"], p["freq"], 0.0)
    # the CPU twin, from the device's (sorted) state
    sim = O.MirSim(nodes, p["dx"], T=T, sort_every=4, seed=seed)
    for sid, (q, m) in enumerate([(-p["q"], p["me"]), (p["q"], p["mi"])]):
        sp = O.MirSpecies(1, cap, q, m, p["w0"], bnd=(2, 2))
        st = ctx.get_state(sid)
        sp.set_state(st["cell"], st["ox"], None, st["vx"], st["vy"], st["vz"], st["id"])
        sim.add_species(sp)
    sim.add_mcc(0, 1, [0, 2, 3], xs["energies"], xs["electron"], p["mi"], 0.0, p["Tg"], tn)
    sim.add_mcc(1, None, [1], [0.0], xs["ion"], p["mi"], 0.0, p["Tg"], tn)
    sim.set_poisson(p["eps0"], tags, 0)
    sim.wf = {1: (0.0, p["volts"], p["freq"], 0.0)}
    return ctx, sim, p


def _assert_species_equal(ctx, sim, sid, what):
    st = ctx.get_state(sid)
    sp = sim.species[sid]
    assert len(st["cell"]) == sp.np, (what, len(st["cell"]), sp.np)
    n = sp.np
    for k in ("cell", "ox", "vx", "vy", "vz", "id"):
        a = st[k]
        b = getattr(sp, k)[:n]
        if not np.array_equal(a.view(np.uint32), b.view(np.uint32)):
            bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
            raise AssertionError(f"{what}: species {sid} field {k} differs at {bad[:5]} "
                                 f"(gpu {a[bad[:3]]} vs cpu {b[bad[:3]]})")


def test_mcc_single_step_bitexact(gpu_ready, orc):
    ctx, sim, p = _ccp_setup()
    dt = p["dt"]
    ev_e = ctx.mcc_run(0, dt)
    ev_i = ctx.mcc_run(1, dt)
    ce = sim.mcc_run(0, dt)
    ci = sim.mcc_run(1, dt)
    assert np.array_equal(ev_e[:3], ce) and np.array_equal(ev_i[:1], ci)
    assert ev_e[0] > 0 and ev_e[2] > 0, ev_e  # elastic and ionisation happened
    _assert_species_equal(ctx, sim, 0, "after mcc")
    _assert_species_equal(ctx, sim, 1, "after mcc")
    nu = ctx.mcc_counts(0, 3)
    assert np.array_equal(nu.sum(axis=1), ev_e[:3])


def test_ccp_loop_bitexact(gpu_ready, orc):
    ctx, sim, p = _ccp_setup(ne=8000, ni=8000)
    dt = p["dt"]
    for chunk in range(3):
        ctx.step(dt, 10)
        sim.step(dt, 10)
        _assert_species_equal(ctx, sim, 0, f"after {10 * (chunk + 1)} steps")
        _assert_species_equal(ctx, sim, 1, f"after {10 * (chunk + 1)} steps")
        from iskra_amd import _lib as L
        assert np.array_equal(ctx.field(L.F_PHI), sim.phi)
        assert np.array_equal(ctx.field(L.F_EX), sim.E.astype(np.float64))
    assert sim.mccs[0]["events"][2] > 0  # ionisation occurred during the run


def test_two_stream_loop_bitexact(gpu_ready, orc):
    from iskra_amd import Context, _lib as L, problems as P
    import oracle as O
    p = P.two_stream_1d()
    ctx = Context()
    ctx.grid(1, (p["nodes"],), (p["dx"],))
    ctx.tuning(sort_every=4)
    n = len(p["x"])
    e = ctx.species(p["q"], p["m"], p["w0"], n, track_ids=True)
    ctx.boundary(e, [L.BND_WRAP])
    ctx.upload(e, p["x"], p["v"])
    ctx.set_field(L.F_BG, np.full(p["nodes"], p["bg"]))
    ctx.poisson(L.POISSON_DIRECT, p["eps0"], None, periodic_x=True)
    sim = O.MirSim(p["nodes"], p["dx"], T=64, sort_every=4)
    sp = O.MirSpecies(1, n, p["q"], p["m"], p["w0"], bnd=(1, 1))
    st = ctx.get_state(e)
    sp.set_state(st["cell"], st["ox"], None, st["vx"], st["vy"], st["vz"], st["id"])
    sim.add_species(sp)
    sim.bg[:] = p["bg"]
    sim.set_poisson(p["eps0"], np.zeros(p["nodes"], np.int8), 1)
    for k in range(4):
        ctx.step(p["dt"], 25)
        sim.step(p["dt"], 25)
        _assert_species_equal(ctx, sim, 0, f"two-stream after {25 * (k + 1)} steps")
        assert np.array_equal(ctx.field(L.F_PHI), sim.phi)
        assert np.array_equal(ctx.field(L.F_RHO), sim.rho)
    s = ctx.stats()
    assert s["steps"] == 100 and s["particle_steps"] == 100 * n


def test_slab_2d_loop_bitexact(gpu_ready, orc):
    """The 2-D time loop over 40 steps against the CPU twin: each step the
    device's rho must equal the twin's bit for bit; the device's potential
    (spectral solve) must match the FP64 banded solve of that rho to 1e-12
    and is then handed to the twin, whose E = -grad phi must equal the
    device's bit for bit; every 8 steps the particle states (electrons
    sorted every 4 steps, ions every 16) must be equal bit for bit."""
    from iskra_amd import Context, _lib as L, problems as P
    p = P.slab_2d(cells=64, ppc=8)
    nodes, dx, n = p["nodes"], p["dx"], p["n_part"]
    ctx = Context()
    ctx.grid(2, (nodes, nodes), (dx, dx))
    ctx.tuning(tile=8, chunk=1024, margin=2, sort_every=4)
    tags = np.zeros((nodes, nodes), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    ctx.poisson(L.POISSON_SPECTRAL, p["eps0"], tags.ravel(), rtol=1e-10)
    ctx.tag_value(1, 0.0)
    sim = O_MirSim(orc, nodes, dx)
    sim.set_poisson_2d(p["eps0"], tags, [0.0, 0.0])
    sids = []
    try:
        for k, (q, m, vth) in enumerate([(-p["q"], p["me"], p["ve"]), (p["q"], p["mi"], p["vi"])]):
            s = ctx.species(q, m, p["w0"], n, track_ids=True)
            ctx.boundary(s, [L.BND_DISCARD, L.BND_DISCARD])
            ctx.sample_maxwellian(s, n, [p["L"]] * 2, [vth] * 3, seed=11, stream=k)
            st = ctx.get_state(s)
            sp = orc.MirSpecies(2, n, q, m, p["w0"], bnd=(2, 2))
            sp.set_state(st["cell"], st["ox"], st["oy"], st["vx"], st["vy"], st["vz"], st["id"])
            sim.add_species(sp)
            sids.append(s)
        ctx.species_sort_every(sids[1], 16)
        sim.species[1].sort_every = 16
        for it in range(40):
            ctx.step(p["dt"], 1)
            sim.advance(p["dt"])
            rho = ctx.field(L.F_RHO)
            assert np.array_equal(rho.view(np.uint64), sim.rho.view(np.uint64)), it
            phi = ctx.field(L.F_PHI)
            ref = orc.potential_banded(nodes, nodes, dx, tags, [0.0, 0.0], rho, p["eps0"])
            assert np.max(np.abs(phi - ref)) <= 1e-12 * np.max(np.abs(ref)), it
            sim.set_phi(phi)
            assert np.array_equal(ctx.field(L.F_EX), sim.E[0::2].astype(np.float64)), it
            assert np.array_equal(ctx.field(L.F_EY), sim.E[1::2].astype(np.float64)), it
            sim.end_step(p["dt"])
            if it % 8 == 7:
                for k in range(2):
                    _assert_species_equal_2d(ctx, sim, k, f"2-D step {it + 1}")
    finally:
        ctx.close()


def O_MirSim(orc, nodes, dx):
    return orc.MirSim(nodes, dx, T=8, sort_every=4, ny=nodes, dy=dx)


def _assert_species_equal_2d(ctx, sim, sid, what):
    st = ctx.get_state(sid)
    sp = sim.species[sid]
    assert len(st["cell"]) == sp.np, (what, len(st["cell"]), sp.np)
    for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id"):
        a, b = st[k], getattr(sp, k)[:sp.np]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (what, sid, k)
