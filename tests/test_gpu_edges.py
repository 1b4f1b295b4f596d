"""GPU edge cases of the step: species that are empty or emptied by the walls,
particle counts that fill no whole work-group or chunk, and the MCC limits
(probability assertion, particle capacity exhausted by ionisation products).

The deposits are compared bit for bit with the oracle's mirror of the
fused push (the same fixed-point arithmetic), as in test_gpu_parity.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIX = 2.0 ** -32


def _state(rng, n, ncx, ncy, vx_scale, vy_scale):
    i = rng.integers(0, ncx, n)
    j = rng.integers(0, ncy, n)
    cell = (i | (j << 16)).astype(np.int32)
    ox = rng.random(n, dtype=np.float32)
    oy = rng.random(n, dtype=np.float32)
    vx = (rng.standard_normal(n) * vx_scale).astype(np.float32)
    vy = (rng.standard_normal(n) * vy_scale).astype(np.float32)
    vz = rng.standard_normal(n).astype(np.float32)
    ids = rng.permutation(n).astype(np.uint32)
    return cell, ox, oy, vx, vy, vz, ids


def _slab(nodes=33, solver=None):
    from iskra_amd import Context, _lib as L
    ctx = Context()
    dx = 1e-4
    ctx.grid(2, (nodes, nodes), (dx, dx))
    tags = np.zeros((nodes, nodes), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    ctx.poisson(L.POISSON_SPECTRAL if solver is None else solver, 8.854e-12, tags.ravel(),
                rtol=1e-8)
    ctx.tag_value(1, 0.0)
    return ctx, dx


def test_empty_species_step(gpu_ready):
    """Two species with no particles: the loop runs, rho and phi stay zero."""
    from iskra_amd import _lib as L
    ctx, dx = _slab()
    e = ctx.species(-1.6e-19, 9.1e-31, 1e6, 1000)
    i = ctx.species(1.6e-19, 6.6e-26, 1e6, 1000)
    ctx.step(1e-12, 3)
    ctx.sync()
    assert ctx.count(e) == 0 and ctx.count(i) == 0
    assert np.all(ctx.field(L.F_RHO) == 0.0)
    assert np.all(ctx.field(L.F_PHI) == 0.0)
    s = ctx.stats()
    assert s["steps"] == 3 and s["particle_steps"] == 0
    ctx.close()


def test_species_emptied_by_the_walls(gpu_ready):
    """Every particle leaves through a discarding wall in the first step; the
    loop goes on with an empty species (and a sort of nothing)."""
    from iskra_amd import _lib as L
    ctx, dx = _slab()
    ctx.tuning(sort_every=2)
    rng = np.random.default_rng(5)
    n = 3000
    dt = 1e-12
    sid = ctx.species(-1.6e-19, 9.1e-31, 1e6, n)
    ctx.boundary(sid, [L.BND_DISCARD, L.BND_DISCARD])
    cell, ox, oy, vx, vy, vz, ids = _state(rng, n, 32, 32, 0.0, 0.0)
    vx[:] = np.float32(100.0 * dx / dt)  # 100 cells per step, out of a 32-cell grid
    ctx.set_state(sid, cell, ox, oy, vx, vy, vz, ids)
    ctx.step(dt, 1)
    ctx.sync()
    assert ctx.count(sid) == 0
    ctx.step(dt, 4)
    ctx.sync()
    assert ctx.count(sid) == 0
    assert np.all(ctx.field(L.F_RHO) == 0.0)
    ctx.close()


@pytest.mark.parametrize("n", [1, 63, 2049, 4099])
def test_ragged_counts_push_deposit_bitexact(gpu_ready, orc, n):
    """Counts that fill no whole work-group or chunk (1, 63, 2049, 4099): the
    push and the fixed-point deposit equal the oracle's mirror bit for bit."""
    from iskra_amd import Context, _lib as L
    rng = np.random.default_rng(n)
    nx, ny = 41, 37
    dx, dy = 1e-3, 1.1e-3
    q, m, w0, dt = -1.6e-19, 9.1e-31, 1e6, 1e-11
    ctx = Context()
    ctx.grid(2, (nx, ny), (dx, dy))
    sid = ctx.species(q, m, w0, n, track_ids=True)
    ctx.boundary(sid, [L.BND_DISCARD, L.BND_WRAP])
    st = _state(rng, n, nx - 1, ny - 1, 0.7 * dx / dt, 0.7 * dy / dt)
    ctx.set_state(sid, *st)
    Ex = rng.standard_normal(nx * ny) * 3e4
    Ey = rng.standard_normal(nx * ny) * 3e4
    ctx.set_field(L.F_EX, Ex)
    ctx.set_field(L.F_EY, Ey)
    E32 = np.empty(2 * nx * ny, dtype=np.float32)
    E32[0::2] = Ex.astype(np.float32)
    E32[1::2] = Ey.astype(np.float32)
    V = orc.cell_volume(2, nx, ny, dx, dy)
    st0 = ctx.get_state(sid)
    ctx.push(sid, dt)
    st1 = ctx.get_state(sid)
    dens = ctx.field(L.F_DENSITY, sid)
    ms = {k: v.copy() for k, v in st0.items()}
    cnt = np.zeros(nx * ny, dtype=np.int64)
    live = orc.lib().mir_push_deposit(
        2, n, orc.ptr(ms["cell"]), orc.ptr(ms["ox"]), orc.ptr(ms["oy"]), orc.ptr(ms["vx"]),
        orc.ptr(ms["vy"]), orc.ptr(E32), nx, nx - 1, ny - 1, np.float32(q / m * dt),
        np.float32(dt / dx), np.float32(dt / dy), L.BND_DISCARD, L.BND_WRAP, 1, orc.ptr(cnt))
    for k in ("cell", "ox", "oy", "vx", "vy"):
        assert np.array_equal(st1[k].view(np.uint32), ms[k].view(np.uint32)), k
    assert np.array_equal(dens, (cnt.astype(np.float64) * FIX) * w0 / V)
    assert ctx.count(sid) == live
    ctx.close()


def _mcc_1d(ne, cap, types, energies, tables, ng, Te_eV, seed=3):
    from iskra_amd import Context, _lib as L, problems as P
    nodes = 65
    ctx = Context()
    ctx.grid(1, (nodes,), (1e-4,))
    ctx.rng(seed)
    e = ctx.species(-P.QE, P.ME, 1e6, cap, track_ids=True)
    i = ctx.species(P.QE, 6.6e-26, 1e6, cap, track_ids=True)
    rng = np.random.default_rng(seed)
    ve = P.thermal_speed(Te_eV * P.QE / P.KB, P.ME)
    ctx.upload(e, (rng.random(ne) * 64e-4)[:, None], rng.standard_normal((ne, 3)) * ve)
    mid = ctx.mcc(e, types, energies, tables, ion=i, target_m=6.6e-26, target_q=0.0,
                  target_T=300.0)
    ctx.set_field(L.F_TARGET, np.full(nodes, ng), index=mid)
    return ctx, e, i, mid


def test_mcc_probability_assertion(gpu_ready):
    """max_P above 1/N (mcc.jl:244) is refused before any kernel runs."""
    from iskra_amd import _lib as L
    flat = (np.array([0.0, 1e4]), np.array([1e-18, 1e-18]))
    ctx, e, i, mid = _mcc_1d(1000, 2000, [L.COLL_ELASTIC_ISOTROPIC, L.COLL_EXCITATION,
                                          L.COLL_IONIZATION], [0.0, 11.5, 15.8],
                             [flat, flat, flat], ng=1e23, Te_eV=50.0)
    with pytest.raises(L.IskraError) as ex:
        ctx.mcc_run(mid, 1e-11)
    assert ex.value.code == L.E_PROB
    ctx.close()


def test_ionisation_beyond_capacity(gpu_ready):
    """Ionisation products that do not fit: both species stop at their
    capacity, no write goes past it, and the call (iskra_mcc_run ends with
    iskra_sync) reports the overflow; so does every later sync."""
    from iskra_amd import _lib as L
    ne, cap = 2000, 2100
    flat = (np.array([0.0, 1e4]), np.array([1e-18, 1e-18]))
    ctx, e, i, mid = _mcc_1d(ne, cap, [L.COLL_IONIZATION], [15.8], [flat], ng=1e22,
                             Te_eV=50.0)
    with pytest.raises(L.IskraError) as ex:
        ctx.mcc_run(mid, 1e-11)
    assert ex.value.code == L.E_CAPACITY
    with pytest.raises(L.IskraError) as ex:
        ctx.sync()
    assert ex.value.code == L.E_CAPACITY
    nu = ctx.mcc_counts(mid, 1)
    assert nu.sum() > cap - ne  # more ionisations than free slots
    # electrons: ne + products clipped to the capacity; ions (none loaded):
    # one per product, fewer than the capacity here
    assert ctx.count(e) == cap
    assert 0 < ctx.count(i) <= cap
    st = ctx.get_state(e)
    assert len(st["cell"]) == cap and np.all((st["cell"] >= 0) & (st["cell"] < 64))
    ctx.close()


@pytest.mark.parametrize("nx,ny", [(17, 17), (33, 9)])
def test_singular_periodic_pcg(gpu_ready, orc, nx, ny):
    """A fully periodic 2-D system has no Dirichlet node and is singular
    (generalized_poisson.jl:202; the reference's A\\b fixes the constant by
    round-off).  The solve projects the right-hand side to zero mean and
    returns the mean-free solution: it must equal the minimum-norm
    least-squares solution of the reference matrix (assembled with its
    periodic-ring quirks) for the projected density, which is mean-free
    because the constants span the null space."""
    from iskra_amd import Context, _lib as L
    rng = np.random.default_rng(nx * ny)
    dx = dy = 0.01
    eps0 = 8.854e-12
    ctx = Context()
    ctx.grid(2, (nx, ny), (dx, dy))
    ctx.poisson(L.POISSON_PCG, eps0, None, periodic_x=True, periodic_y=True, rtol=1e-10)
    rho = rng.standard_normal(nx * ny) * 1e-9 + 3e-10  # nonzero mean on purpose
    ctx.set_field(L.F_RHO, rho)
    it, rr = ctx.solve()
    assert rr <= 1e-10
    phi = ctx.field(L.F_PHI)
    A, b, isdof = orc.poisson_matrix(2, nx, ny, dx, dy, periodic=[1, 2])
    assert np.all(isdof == 1)
    bb = -(rho - rho.mean()) / eps0
    x = np.linalg.lstsq(A.reshape(nx * ny, nx * ny), bb, rcond=None)[0]
    assert abs(x.mean()) <= 1e-10 * np.abs(x).max()
    assert abs(phi.mean()) <= 1e-9 * np.abs(x).max()
    assert np.max(np.abs(phi - x)) <= 1e-7 * np.abs(x).max()
    ctx.close()


def test_source_box_outside_the_grid_is_refused(gpu_ready):
    """sample!(::MaxwellianSource) draws x = U wx + dx (sources.jl:24-34); a
    box reaching past the grid has no cell/offset representation and is
    refused with ISKRA_E_ARG before anything is drawn; a box of exactly the
    domain stays inside it."""
    from iskra_amd import Context, _lib as L
    ctx = Context()
    try:
        ncx, L_ = 300, 0.025
        dx = L_ / ncx
        ctx.grid(1, (ncx + 1,), (dx,))
        s = ctx.species(-1.0, 1.0, 1.0, 10000, track_ids=True)
        for wx, x0 in (([L_ * 1.01], [0.0]), ([L_ / 2], [L_ * 0.6]), ([L_], [-dx])):
            with pytest.raises(L.IskraError) as ei:
                ctx.sample_maxwellian(s, 100, wx, [1.0, 1.0, 1.0], dx=x0, seed=1, stream=0)
            assert ei.value.code == L.E_ARG
            assert ctx.count(s) == 0
        ctx.sample_maxwellian(s, 5000, [L_], [1.0, 1.0, 1.0], seed=1, stream=0)
        st = ctx.get_state(s)
        assert len(st["cell"]) == 5000
        assert st["cell"].min() >= 0 and st["cell"].max() <= ncx - 1
        assert st["ox"].min() >= 0.0 and st["ox"].max() < 1.0
    finally:
        ctx.close()
