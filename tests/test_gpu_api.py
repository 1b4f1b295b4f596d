"""The reference-named API end to end on the GPU: the reference's own FDM
test, and whole `solve` runs written like the reference's problem scripts,
checked bit for bit against the same runs driven through the C ABI directly
(whose parity with the oracle is covered in test_gpu_loop.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def pic_module():
    import iskra_amd.particle_in_cell as pic
    saved = (pic.enter_loop, pic.after_loop, pic.exit_loop, pic.after_push)
    yield pic
    pic.enter_loop, pic.after_loop, pic.exit_loop, pic.after_push = saved


def test_parallel_capacitor_like_reference(gpu_ready):
    # FiniteDifferenceMethod/test/runtests.jl:8-22, through the mirrored names
    from iskra_amd import finite_difference_method as fdm
    from iskra_amd import regular_grids as rg
    grid = rg.create_uniform_grid(np.arange(0, 1.01, 0.5), np.arange(0, 2.01, 0.5))
    solv = fdm.create_poisson_solver(grid, 1.0, rtol=1e-12)
    bcs = np.zeros(grid.size(), np.int8)
    bcs[0, 0:5] = 1
    bcs[2, 0:5] = 2
    fdm.apply_dirichlet(solv, bcs == 1, 0)
    fdm.apply_dirichlet(solv, bcs == 2, 1)
    phi = fdm.calculate_electric_potential(solv, np.zeros(grid.size()))
    E = fdm.calculate_electric_field(solv, phi)
    # E is stored in FP32 on the device: tolerance 1e-6
    assert np.allclose(E[:, :, 0], -np.ones((3, 5)), atol=1e-6)
    assert np.allclose(E[:, :, 1], np.zeros((3, 5)), atol=1e-6)
    assert np.allclose(phi, np.repeat([[0.0], [0.5], [1.0]], 5, axis=1), atol=1e-10)


def _two_stream_state():
    from iskra_amd import problems as P
    return P.two_stream_1d(particles=4000)


def test_two_stream_solve_matches_c_abi(gpu_ready, pic_module):
    pic = pic_module
    from iskra_amd import Context, _lib as L
    from iskra_amd import finite_difference_method as fdm
    from iskra_amd import regular_grids as rg
    p = _two_stream_state()
    n = len(p["x"])
    steps = 40

    # reference-style script (problem/ts_unstable.jl structure)
    xs = np.arange(p["nodes"]) * p["dx"]
    grid = rg.create_uniform_grid(xs)
    solver = fdm.create_poisson_solver(grid, p["eps0"])
    fdm.apply_periodic(solver)
    e = pic.create_kinetic_species("e-", n, p["q"], p["m"], p["w0"], D=1)
    ions = pic.FluidSpecies("i+", 0.0, 1.0, 1.0, np.full(p["nodes"], p["bg"]), 0.0)
    config = pic.Config(grid=grid, solver=solver, pusher=pic.create_boris_pusher(),
                        species=[e, ions])

    def start():
        e.x[:n, 0] = p["x"][:, 0]
        e.v[:n] = p["v"]
        e.np = n
    pic.enter_loop = start
    pic.after_push = lambda part, g: pic.wrap(part, g)
    pic.solve(config, p["dt"], steps)
    phi_api = pic.field(config, "phi")
    st_api = config.ctx.get_state(e._sid)

    # the same run through the C ABI
    ctx = Context(0)
    ctx.grid(1, (p["nodes"],), (p["dx"],))
    ctx.rng(0)
    sid = ctx.species(p["q"], p["m"], p["w0"], n, track_ids=True)
    ctx.boundary(sid, [L.BND_WRAP])
    ctx.upload(sid, p["x"], p["v"], np.arange(1, n + 1, dtype=np.uint32))
    ctx.set_field(L.F_BG, np.full(p["nodes"], p["bg"]))
    ctx.poisson(L.POISSON_DIRECT, p["eps0"], None, periodic_x=True)
    ctx.step(p["dt"], steps)
    st = ctx.get_state(sid)
    for k in ("cell", "ox", "vx", "id"):
        assert np.array_equal(st_api[k], st[k]), k
    assert np.array_equal(phi_api, ctx.field(L.F_PHI))
    # host mirror reads back the device particles
    assert e.np == n
    X = e.x[:n, 0]
    assert np.all((X >= 0) & (X < p["L"]))
    ctx.close()


def test_rf_discharge_solve_with_mcc_and_electrode_hook(gpu_ready, pic_module):
    """problem/11_rf_discharge.jl structure: two kinetic species, a fluid
    gas target, electron and ion MCC, discard walls, and the electrode
    voltage set by apply_dirichlet in the per-iteration hook."""
    pic = pic_module
    from iskra_amd import Context, _lib as L, problems as P
    from iskra_amd import chemistry as ch
    from iskra_amd import finite_difference_method as fdm
    from iskra_amd import regular_grids as rg
    nodes, ne, seed, steps = 129, 20000, 5, 25
    p = P.ccp_1d(cells=nodes - 1, electrons=ne, ions=ne)
    rng = np.random.default_rng(seed)
    ve = P.thermal_speed(12.0 * P.QE / P.KB, P.ME)
    xe, xi = rng.random(ne) * p["L"], rng.random(ne) * p["L"]
    vve, vvi = rng.standard_normal((ne, 3)) * ve, rng.standard_normal((ne, 3)) * p["vi"]
    xs = p["xs"]
    cap = 3 * ne

    grid = rg.create_uniform_grid(np.arange(nodes) * p["dx"])
    solver = fdm.create_poisson_solver(grid, p["eps0"])
    bcs = np.zeros(nodes, np.int8)
    bcs[0], bcs[-1] = 1, 2
    fdm.apply_dirichlet(solver, bcs == 1, 0.0)
    fdm.apply_dirichlet(solver, bcs == 2, 0.0)
    e = pic.create_kinetic_species("e-", cap, -p["q"], p["me"], p["w0"], D=1)
    ion = pic.create_kinetic_species("Ar+", cap, p["q"], p["mi"], p["w0"], D=1)
    gas = pic.FluidSpecies("Ar", 1.0, 0.0, p["mi"], np.full(nodes, p["ng"]), p["Tg"])
    sig = [ch.CrossSection(a, b) for a, b in xs["electron"]]
    electron = ch.mcc(ch.reactions([
        (sig[0], [e, gas], [e, gas]),
        (sig[1], [e, gas], [e, gas], ch.Excitation(11.5)),
        (sig[2], [e, gas], [e, e, ion], ch.Ionization(15.8)),
    ]))
    ionmcc = ch.mcc(ch.reactions([
        (ch.CrossSection(*xs["ion"][0]), [ion, gas], [ion, gas], ch.ElasticBackward()),
    ]))
    config = pic.Config(grid=grid, solver=solver, species=[e, ion, gas],
                        interactions=[electron, ionmcc])
    config.seed = seed

    def start():
        e.x[:ne, 0], e.v[:ne], e.np = xe, vve, ne
        ion.x[:ne, 0], ion.v[:ne], ion.np = xi, vvi, ne

    volts = []

    def iteration(i, t, dt):
        v = p["volts"] * np.sin(2 * np.pi * p["freq"] * t)
        volts.append(v)
        fdm.apply_dirichlet(solver, bcs == 1, v)
    pic.enter_loop = start
    pic.after_loop = iteration
    pic.after_push = lambda part, g: pic.discard(part, g, dims=1)
    pic.solve(config, p["dt"], steps)
    api = [config.ctx.get_state(s._sid) for s in (e, ion)]
    phi_api = pic.field(config, "phi")
    nu_api = electron.counts()

    ctx = Context(0)
    ctx.grid(1, (nodes,), (p["dx"],))
    ctx.rng(seed)
    se = ctx.species(-p["q"], p["me"], p["w0"], cap, track_ids=True)
    si = ctx.species(p["q"], p["mi"], p["w0"], cap, track_ids=True)
    ctx.boundary(se, [L.BND_DISCARD])
    ctx.boundary(si, [L.BND_DISCARD])
    ctx.upload(se, xe[:, None], vve, np.arange(1, ne + 1, dtype=np.uint32))
    ctx.upload(si, xi[:, None], vvi, np.arange(1, ne + 1, dtype=np.uint32))
    me_id = ctx.mcc(se, [L.COLL_ELASTIC_ISOTROPIC, L.COLL_EXCITATION, L.COLL_IONIZATION],
                    [0.0, 11.5, 15.8], xs["electron"], ion=si, target_m=p["mi"], target_q=0.0,
                    target_T=p["Tg"])
    mi_id = ctx.mcc(si, [L.COLL_ELASTIC_BACKWARD], [0.0], xs["ion"], ion=-1, target_m=p["mi"],
                    target_q=0.0, target_T=p["Tg"])
    ctx.set_field(L.F_TARGET, np.full(nodes, p["ng"]), index=me_id)
    ctx.set_field(L.F_TARGET, np.full(nodes, p["ng"]), index=mi_id)
    ctx.poisson(L.POISSON_DIRECT, p["eps0"], bcs)
    for it in range(steps):
        ctx.step(p["dt"], 1)
        ctx.tag_value(1, volts[it])
    for sid, sa in zip((se, si), api):
        sb = ctx.get_state(sid)
        for k in ("cell", "ox", "vx", "vy", "vz", "id"):
            assert np.array_equal(sa[k], sb[k]), (sid, k)
    assert np.array_equal(phi_api, ctx.field(L.F_PHI))
    assert np.array_equal(nu_api, ctx.mcc_counts(me_id, 3))
    assert nu_api.sum() > 0
    ctx.close()


def test_density_and_gather_operators(gpu_ready, pic_module):
    """density (kinetic.jl:53), particle_to_grid and grid_to_particle
    (cloud_in_cell.jl) on a bound 2-D configuration against the FP64
    reference restatement in the oracle."""
    pic = pic_module
    import oracle as O
    from iskra_amd import finite_difference_method as fdm
    from iskra_amd import regular_grids as rg
    nx, ny, n = 17, 9, 3000
    dx = 0.25
    grid = rg.create_uniform_grid(np.arange(nx) * dx, np.arange(ny) * dx)
    solver = fdm.create_poisson_solver(grid, 1.0)
    walls = np.zeros((nx, ny), bool)
    walls[0, :] = walls[-1, :] = walls[:, 0] = walls[:, -1] = True
    fdm.apply_dirichlet(solver, walls, 0.0)
    e = pic.create_kinetic_species("e-", n, -1.0, 1.0, 0.5, D=2)
    config = pic.Config(grid=grid, solver=solver, species=[e])
    rng = np.random.default_rng(3)
    x = np.column_stack([rng.random(n) * (nx - 1) * dx, rng.random(n) * (ny - 1) * dx])
    e.x[:n] = x
    e.v[:n] = 0.0
    e.np = n
    pic.bind(config)
    dens = pic.density(e, grid)
    u = O.particle_to_grid(2, x.T.ravel(), np.full(n, 0.5), nx, ny, dx, dx)
    u = u.reshape(ny, nx).T
    V = rg.cell_volume(grid)
    # FP32 in-cell offsets, fixed-point (2^-32) weights: relative 1e-6
    assert np.allclose(dens * V, u, rtol=1e-6, atol=1e-6 * 0.5)
    assert np.isclose((dens * V).sum(), 0.5 * n, rtol=1e-12)
    E = np.zeros((nx, ny, 3))
    E[..., 0] = 2.0 * grid.coords[0] - 0.5 * grid.coords[1]
    E[..., 1] = 1.0 + grid.coords[1]
    Ep = pic.grid_to_particle(grid, e, E)
    # CIC reproduces a linear field exactly up to FP32 offsets and E storage
    st = config.ctx.download(e._sid)
    xp = st[0]
    assert np.allclose(Ep[:, 0], 2.0 * xp[:, 0] - 0.5 * xp[:, 1], atol=1e-5)
    assert np.allclose(Ep[:, 1], 1.0 + xp[:, 1], atol=1e-5)


def test_sort_schedule_does_not_change_results(gpu_ready):
    """The counting sort only orders particles: electrons sorted every 4 steps
    and ions every 4, 16 or 1000 steps (iskra_species_sort_every) give the
    same rho, phi and particle states, bit for bit, compared by id."""
    from iskra_amd import Context, _lib as L, problems as P
    p = P.slab_2d(cells=64, ppc=16)
    n = p["nodes"]
    out = []
    for ion_every in (0, 16, 1000):
        ctx = Context()
        ctx.grid(2, (n, n), (p["dx"], p["dx"]))
        ctx.tuning(sort_every=4)
        tags = np.zeros((n, n), np.int8)
        tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
        ctx.poisson(L.POISSON_SPECTRAL, p["eps0"], tags.ravel(), rtol=1e-10)
        ctx.tag_value(1, 0.0)
        e = ctx.species(-p["q"], p["me"], p["w0"], p["n_part"] + 64, track_ids=True)
        i = ctx.species(p["q"], p["mi"], p["w0"], p["n_part"] + 64, track_ids=True)
        for s in (e, i):
            ctx.boundary(s, [L.BND_DISCARD, L.BND_DISCARD])
        ctx.species_sort_every(i, ion_every)
        ctx.sample_maxwellian(e, p["n_part"], [p["L"]] * 2, [p["ve"]] * 3, seed=3, stream=0)
        ctx.sample_maxwellian(i, p["n_part"], [p["L"]] * 2, [p["vi"]] * 3, seed=3, stream=1)
        ctx.step(p["dt"], 40)
        st = []
        for s in (e, i):
            g = ctx.get_state(s)
            live = g["cell"] >= 0
            o = np.argsort(g["id"][live], kind="stable")
            st.append({k: v[live][o] for k, v in g.items()})
        out.append((ctx.field(L.F_RHO), ctx.field(L.F_PHI), st))
        ctx.close()
    r0, f0, s0 = out[0]
    for r, f, s in out[1:]:
        assert np.array_equal(r.view(np.uint64), r0.view(np.uint64))
        assert np.array_equal(f.view(np.uint64), f0.view(np.uint64))
        for a, b in zip(s, s0):
            for k in a:
                assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def _ccp_run(tile, chunk, sort_every, steps=60, ne=30000):
    """A 1-D RF discharge with MCC (elastic, excitation, ionisation; ion
    backscatter) under the given work decomposition; states by id."""
    from iskra_amd import Context, _lib as L, problems as P
    p = P.ccp_1d(cells=256, electrons=ne, ions=ne)
    nodes = p["nodes"]
    rng = np.random.default_rng(5)
    ctx = Context()
    ctx.grid(1, (nodes,), (p["dx"],))
    ctx.tuning(tile=tile, chunk=chunk, sort_every=sort_every)
    ctx.rng(9)
    e = ctx.species(-p["q"], p["me"], p["w0"], 4 * ne, track_ids=True)
    i = ctx.species(p["q"], p["mi"], p["w0"], 4 * ne, track_ids=True)
    for s in (e, i):
        ctx.boundary(s, [L.BND_DISCARD])
    ve = P.thermal_speed(15.0 * P.QE / P.KB, P.ME)
    ctx.upload(e, rng.random(ne)[:, None] * p["L"], rng.standard_normal((ne, 3)) * ve)
    ctx.upload(i, rng.random(ne)[:, None] * p["L"], rng.standard_normal((ne, 3)) * p["vi"])
    xs = p["xs"]
    me = ctx.mcc(e, [L.COLL_ELASTIC_ISOTROPIC, L.COLL_EXCITATION, L.COLL_IONIZATION],
                 xs["energies"], xs["electron"], ion=i, target_m=p["mi"], target_T=p["Tg"])
    mi = ctx.mcc(i, [L.COLL_ELASTIC_BACKWARD], [0.0], xs["ion"], target_m=p["mi"],
                 target_T=p["Tg"])
    for m in (me, mi):
        ctx.set_field(L.F_TARGET, np.full(nodes, p["ng"]), index=m)
    tags = np.zeros(nodes, np.int8)
    tags[0], tags[-1] = 1, 2
    ctx.poisson(L.POISSON_DIRECT, p["eps0"], tags)
    ctx.tag_waveform(1, 0.0, p["volts"], p["freq"], 0.0)
    ctx.step(p["dt"], steps)
    st = []
    for s in (e, i):
        g = ctx.get_state(s)
        live = g["cell"] >= 0
        o = np.argsort(g["id"][live], kind="stable")
        st.append({k: v[live][o] for k, v in g.items()})
    out = (ctx.field(L.F_RHO), ctx.field(L.F_PHI), st, ctx.mcc_counts(me, 3))
    ctx.close()
    return out


def test_work_decomposition_does_not_change_results_with_ionisation(gpu_ready):
    """Ionisation products are numbered in the order of their parents' ids, so
    neither the sort interval nor the tile and chunk shapes (which change the
    order particles are stored in) change any result: rho, phi, every
    particle's state compared by id, and the collision counts, bit for bit."""
    runs = [_ccp_run(64, 2048, 4), _ccp_run(32, 1024, 16), _ccp_run(128, 4096, 7)]
    r0, f0, s0, n0 = runs[0]
    assert len(s0[0]["id"]) > 30000  # ionisation added electrons
    assert n0[2].sum() > 100
    for r, f, s, nu in runs[1:]:
        assert np.array_equal(r.view(np.uint64), r0.view(np.uint64))
        assert np.array_equal(f.view(np.uint64), f0.view(np.uint64))
        assert np.array_equal(nu, n0)
        for a, b in zip(s, s0):
            assert len(a["id"]) == len(b["id"])
            for k in a:
                assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
