"""Mobile fluid species on the GPU (advance!(::FluidSpecies),
ParticleInCell.jl:74-82, with calculate_advection_diffusion,
FiniteDifferenceMethod/src/advection_diffusion.jl:4-38): the device update
bit for bit against the oracle's FP64 restatement fed the device's own E,
the fluid's charge in rho, and a fluid-only run written like
problem/02_fluid_ions.jl against the reference's FP64 loop."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def pic_module():
    import iskra_amd.particle_in_cell as pic
    saved = (pic.enter_loop, pic.after_loop, pic.exit_loop, pic.after_push)
    yield pic
    pic.enter_loop, pic.after_loop, pic.exit_loop, pic.after_push = saved


def _grid_E(ctx, nodes, dim, rng, scale):
    from iskra_amd import _lib as L
    E = rng.normal(size=(dim,) + (int(np.prod(nodes)),)) * scale
    ctx.set_field(L.F_EX, E[0])
    if dim == 2:
        ctx.set_field(L.F_EY, E[1])


@pytest.mark.parametrize("dim", [1, 2])
def test_fluid_advance_matches_the_restatement_bit_for_bit(gpu_ready, dim):
    import oracle as O
    from iskra_amd import Context, _lib as L
    rng = np.random.default_rng(10 + dim)
    nodes = (37,) if dim == 1 else (23, 17)
    dh = (0.3,) if dim == 1 else (0.3, 0.2)
    nn = int(np.prod(nodes))
    ctx = Context(0)
    ctx.grid(dim, nodes, dh)
    bg_fixed = rng.normal(size=nn)
    ctx.set_field(L.F_BG, bg_fixed)
    fl = [dict(q=1.6, m=3.0, mu=0.05), dict(q=-0.7, m=0.5, mu=0.0)]
    ns = [rng.random(nn) + 0.5 for _ in fl]
    ids = [ctx.fluid(f["q"], f["m"], f["mu"], n) for f, n in zip(fl, ns)]
    assert ids == [0, 1]
    assert np.array_equal(ctx.field(L.F_BG), bg_fixed)  # the fixed part
    dt = 0.01
    for it in range(3):
        _grid_E(ctx, nodes, dim, rng, 2.0)
        Ex = ctx.field(L.F_EX)
        Ey = ctx.field(L.F_EY) if dim == 2 else np.zeros(nn)
        print(f"fluid advance dim {dim} call {it}", flush=True)
        ctx.fluid_advance(dt)
        for k, f in enumerate(fl):
            vx = (Ex * dt) * (f["q"] / f["m"])
            vy = (Ey * dt) * (f["q"] / f["m"])
            # the reference's arrays: n[i(, j)], v[i(, j), component]
            shape = nodes[::-1]
            nJ = ns[k].reshape(shape).T
            v = np.stack([vx.reshape(shape).T, vy.reshape(shape).T, np.zeros(shape).T], axis=-1)
            dn = O.ref_advection_diffusion(nJ, f["mu"], v, dh, dt)
            dn = np.ascontiguousarray(dn.T).ravel()
            ns[k] = ns[k] + dn
            assert np.array_equal(ctx.field(L.F_FLUID_DN, k), dn), (it, k)
            assert np.array_equal(ctx.field(L.F_FLUID_N, k), ns[k]), (it, k)
        # rho with no kinetic species: the fixed background plus q n of each fluid
        ctx.deposit()
        want = (bg_fixed + fl[0]["q"] * ns[0]) + fl[1]["q"] * ns[1]
        assert np.array_equal(ctx.field(L.F_RHO), want), it
    if dim == 1:  # mu = 0 and no advection in 1-D: the second fluid never changes
        assert not ctx.field(L.F_FLUID_DN, 1).any()
    # replacing a fluid's density recomputes the background it contributes
    ns[0] = rng.random(nn)
    ctx.set_field(L.F_FLUID_N, ns[0], index=0)
    ctx.deposit()
    assert np.array_equal(ctx.field(L.F_RHO), (bg_fixed + fl[0]["q"] * ns[0]) + fl[1]["q"] * ns[1])
    ctx.close()


def test_fluid_ions_run_against_the_fp64_loop(gpu_ready, pic_module, tmp_path):
    """problem/02_fluid_ions.jl's set-up on 21 x 21 nodes (5 cm cells, 5 us
    steps, O+ fluid fed by a DensitySource in the start hook) with phi = 0
    walls, 100 steps, against oracle.RefSim (FP64 E and a direct solve).
    The device keeps E in FP32, which the advection term sees:
    densities agree to 1e-9 of their maximum, phi to 1e-6."""
    pic = pic_module
    import oracle as O
    from iskra_amd import finite_difference_method as fdm
    from iskra_amd import problems as P
    from iskra_amd import regular_grids as rg
    nx, dh, dt, steps = 21, 0.05, 5e-6, 100
    L_ = (nx - 1) * dh
    xs = np.arange(nx) * dh
    grid = rg.create_uniform_grid(xs, xs)
    solver = fdm.create_poisson_solver(grid, P.EPS0)
    walls = np.zeros((nx, nx), bool)
    walls[0, :] = walls[-1, :] = walls[:, 0] = walls[:, -1] = True
    fdm.apply_dirichlet(solver, walls, 0.0)
    iO = pic.create_fluid_species("O+", 1.0, P.QE, 8 * 1.67262192e-27, nx, nx)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    delta = 1e5 * np.exp(-(X - 0.5 * L_) ** 2 / (0.03 * L_) - (Y - 0.5 * L_) ** 2 / (0.03 * L_))

    def start():
        iO.n[...] = 0.0
        pic.init(pic.DensitySource(delta, grid), iO, dt)
    pic.enter_loop = start
    config = pic.Config(grid=grid, solver=solver, species=[iO])
    print("fluid ions: 21 x 21 nodes, 100 steps", flush=True)
    pic.solve(config, dt, steps)
    n_dev = iO.n.copy()
    phi_dev = pic.field(config, "phi")

    sim = O.RefSim(2, nx, dh, ny=nx, dy=dh, eps0=P.EPS0)
    sim.set_tags(walls.T.ravel().astype(np.int8))
    f = sim.add_fluid(P.QE, 8 * 1.67262192e-27, 1.0, delta.T.ravel())
    sim.step(dt, steps)
    n_ref = sim.fluids[f]["n"].reshape(nx, nx).T
    phi_ref = sim.phi.reshape(nx, nx).T
    assert np.abs(n_ref - delta).max() > 1e-3 * delta.max()  # the fluid did move
    assert np.allclose(n_dev, n_ref, rtol=0, atol=1e-9 * n_ref.max())
    assert np.allclose(phi_dev, phi_ref, rtol=0, atol=1e-6 * np.abs(phi_ref).max())
    # the fluid's records in a file: n, the last step's dn, v (ParticleInCell.jl:80-81,123)
    from iskra_amd import _lib as L
    from iskra_amd import diagnostics as diag
    from test_openpmd import read_back
    path = diag.new_iteration(lambda it: [diag.save_record(it, k) for k in ("nO+", "dO+", "vO+")],
                              str(tmp_path), steps, steps * dt - dt, dt)
    _, data = read_back(path, tmp_path)
    F = f"/data/{steps}/fields"
    assert np.array_equal(data[F + "/nO+"], n_dev.T)
    assert np.array_equal(data[F + "/dO+"], config.ctx.field(L.F_FLUID_DN, 0).reshape(nx, nx))
    assert data[F + "/dO+"].any() and data[F + "/vO+/x"].any()


def test_batches_refuse_mobile_fluids(gpu_ready):
    from iskra_amd import Context, _lib as L
    from iskra_amd._lib import Batch
    ctx = Context(0)
    ctx.grid(1, (17,), (0.1,))
    ctx.poisson(L.POISSON_DIRECT, 1.0, None, periodic_x=True)
    ctx.fluid(1.0, 1.0, 0.1, np.linspace(1.0, 2.0, 17))
    with pytest.raises(L.IskraError) as e:
        Batch([ctx])
    assert e.value.code == L.E_UNSUPPORTED
    ctx.close()
