"""Axial (r, z) geometry on the device: node volumes, the axisymmetric Poisson
solve against the reference matrix, and the RZ time loop against the CPU
twin bit for bit (RegularGrids.jl:40-53,84-97; generalized_poisson.jl:70-199;
pushers.jl:13-17,52-66; cloud_in_cell.jl:38-73)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _axial_ctx(nr, nz, dr, dz, volper_z=(0, 0)):
    from iskra_amd import Context
    ctx = Context()
    ctx.grid_axial((nr, nz), (dr, dz), vol_periodic_z=volper_z)
    return ctx


@pytest.mark.parametrize("volper_z", [(0, 0), (1, 1)])
def test_axial_density_uses_annulus_volumes(gpu_ready, orc, volper_z):
    """One macro-particle per node position deposits its weight on that node
    only; the density is weight / V with V the reference's annulus volumes."""
    from iskra_amd import _lib as L
    nr, nz, dr, dz = 9, 5, 2.0 ** -9, 2.0 ** -10  # node positions exact in binary
    ctx = _axial_ctx(nr, nz, dr, dz, volper_z)
    try:
        s = ctx.species(-1.0, 1.0, 3.0, 64, track_ids=True)
        i, j = np.meshgrid(np.arange(nr - 1), np.arange(nz - 1), indexing="xy")
        x = np.column_stack([i.ravel() * dr, j.ravel() * dz])
        ctx.upload(s, x, np.zeros((len(x), 3)))
        ctx.deposit_species(s)
        n = ctx.field(L.F_DENSITY, s)
        V = orc.cell_volume_axial(nr, nz, dr, dz, volper_z)
        want = np.zeros(nr * nz)
        k = (j.ravel() * nr + i.ravel())
        want[k] = 3.0 / V[k]
        assert np.allclose(n, want, rtol=1e-15, atol=0)
    finally:
        ctx.close()


def test_axial_pcg_against_reference_lu(gpu_ready, orc):
    """PCG on the weighted axisymmetric system equals the dense LU solve of the
    reference's (non-symmetric) matrix with its Dirichlet rows."""
    from iskra_amd import _lib as L
    nr, nz, dr, dz = 13, 9, 1.5e-3, 1e-3
    eps0 = 8.8541878128e-12
    tags = np.zeros((nz, nr), np.int8)
    tags[:, -1] = 1    # outer wall
    tags[0, :] = 2     # bottom electrode
    tags[-1, :3] = 3   # a small disc on the top edge at the axis
    vals = {1: 0.0, 2: 150.0, 3: -40.0}
    rng = np.random.default_rng(5)
    rho = rng.normal(0.0, 1e-6, nr * nz)
    ctx = _axial_ctx(nr, nz, dr, dz)
    try:
        ctx.poisson(L.POISSON_PCG, eps0, tags.ravel(), rtol=1e-13, max_iter=5000)
        for t, v in vals.items():
            ctx.tag_value(t, v)
        ctx.set_field(L.F_RHO, rho)
        it, rr = ctx.solve(L.WARM_ZERO)
        assert rr <= 1e-13, (it, rr)
        phi = ctx.field(L.F_PHI)
        A, b, isdof = orc.poisson_matrix_rz(nr, nz, dr, dz, dirichlet=[
            ((tags == t).T, v) for t, v in vals.items()])
        ref, sing = orc.potential(A, b, isdof, rho, eps0)
        assert not sing
        scale = np.max(np.abs(ref))
        # rtol 1e-13 bounds the error by cond * 1e-13 (cond ~ 1e3 on this grid)
        assert np.max(np.abs(phi - ref)) <= 1e-9 * scale, np.max(np.abs(phi - ref)) / scale
        assert np.array_equal(phi[tags.ravel() == 2], np.full((tags == 2).sum(), 150.0))
    finally:
        ctx.close()


def test_axial_solver_refusals(gpu_ready):
    from iskra_amd import _lib as L
    nr, nz = 17, 17
    tags = np.zeros((nz, nr), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, -1] = tags[:, 0] = 1
    ctx = _axial_ctx(nr, nz, 1e-3, 1e-3)
    try:
        for kind in (L.POISSON_MG, L.POISSON_MGPCG, L.POISSON_SPECTRAL):
            with pytest.raises(L.IskraError) as ei:
                ctx.poisson(kind, 1.0, tags.ravel())
            assert ei.value.code == L.E_UNSUPPORTED
        with pytest.raises(L.IskraError) as ei:
            ctx.poisson(L.POISSON_PCG, 1.0, tags.ravel(), periodic_y=1)
        assert ei.value.code == L.E_UNSUPPORTED
        with pytest.raises(L.IskraError) as ei:
            ctx.poisson(L.POISSON_PCG, 1.0, np.zeros(nr * nz, np.int8))
        assert ei.value.code == L.E_UNSUPPORTED
    finally:
        ctx.close()


def _assert_species_equal(ctx, sim, sid, what):
    st = ctx.get_state(sid)
    sp = sim.species[sid]
    assert len(st["cell"]) == sp.np, (what, len(st["cell"]), sp.np)
    for k in ("cell", "ox", "oy", "vx", "vy", "vz", "id"):
        a, b = st[k], getattr(sp, k)[:sp.np]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (what, sid, k)


def test_axial_loop_bitexact(gpu_ready, orc):
    """A plasma cylinder (electrons and ions, discarded at the outer wall and
    the z ends, phi = 0 there) over 32 steps of iskra_step with the RZ pusher
    against the CPU twin: each step rho equal bit for bit, the device's PCG
    potential within 1e-8 of the sparse LU solve of the reference matrix and
    then handed to the twin, E equal bit for bit; particle states (sorted
    every 4 steps) equal bit for bit every 8 steps.  The v_theta the RZ
    transform turns is checked to be exercised: nearly every electron's
    v_theta has changed, and the run starts particles within half a cell of
    the axis."""
    from iskra_amd import Context, _lib as L, problems as P
    p = P.slab_2d(cells=32, ppc=8)
    nodes, dx, n = p["nodes"], p["dx"], p["n_part"]
    R = p["L"]
    w0 = 1e16 * np.pi * R * R * R / n
    ctx = Context()
    ctx.grid_axial((nodes, nodes), (dx, dx))
    ctx.tuning(tile=8, chunk=1024, margin=2, sort_every=4)
    tags = np.zeros((nodes, nodes), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, -1] = 1
    ctx.poisson(L.POISSON_PCG, p["eps0"], tags.ravel(), rtol=1e-13, max_iter=20000)
    ctx.tag_value(1, 0.0)
    sim = orc.MirSim(nodes, dx, T=8, sort_every=4, ny=nodes, dy=dx, axial=True)
    sim.set_poisson_2d(p["eps0"], tags, [0.0, 0.0])
    try:
        sids = []
        for k, (q, m, vth) in enumerate([(-p["q"], p["me"], p["ve"]), (p["q"], p["mi"], p["vi"])]):
            s = ctx.species(q, m, w0, n, track_ids=True)
            ctx.boundary(s, [L.BND_DISCARD, L.BND_DISCARD])
            ctx.sample_maxwellian(s, n, [R, R], [vth] * 3, seed=21, stream=k)
            st = ctx.get_state(s)
            sp = orc.MirSpecies(2, n, q, m, w0, bnd=(2, 2))
            sp.set_state(st["cell"], st["ox"], st["oy"], st["vx"], st["vy"], st["vz"], st["id"])
            sim.add_species(sp)
            sids.append(s)
        vz_by_id = np.zeros(n, np.float32)
        vz_by_id[sim.species[0].id[:n]] = sim.species[0].vz[:n]
        r0 = (sim.species[0].cell[:n] & 0xFFFF) + sim.species[0].ox[:n].astype(np.float64)
        for it in range(32):
            ctx.step(p["dt"], 1)
            sim.advance(p["dt"])
            rho = ctx.field(L.F_RHO)
            assert np.array_equal(rho.view(np.uint64), sim.rho.view(np.uint64)), it
            phi = ctx.field(L.F_PHI)
            ref = orc.potential_rz(nodes, nodes, dx, dx, tags, [0.0, 0.0], rho, p["eps0"])
            assert np.max(np.abs(phi - ref)) <= 1e-8 * np.max(np.abs(ref)), it
            sim.set_phi(phi)
            assert np.array_equal(ctx.field(L.F_EX), sim.E[0::2].astype(np.float64)), it
            assert np.array_equal(ctx.field(L.F_EY), sim.E[1::2].astype(np.float64)), it
            sim.end_step(p["dt"])
            if it % 8 == 7:
                for k in range(2):
                    _assert_species_equal(ctx, sim, k, f"RZ step {it + 1}")
        e = sim.species[0]
        assert e.np > n // 2
        # v_theta turned (it is constant under the Cartesian push with B = 0)
        ids = e.id[:e.np]
        assert np.mean(e.vz[:e.np] != vz_by_id[ids]) > 0.9
        assert np.any(r0 < 0.5)
    finally:
        ctx.close()
