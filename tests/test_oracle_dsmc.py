"""The DSMC twin (mir_dsmc: PIC.perform!(::DirectSimulationMonteCarlo),
Chemistry/src/dsmc.jl:81-142, with the elastic collision of dsmc.jl:32-79)
against closed forms: the trial count per cell with its carried remainder,
conservation of momentum and kinetic energy by every collision, the
isotropic post-collision relative velocity, and the index redraw rule."""
import numpy as np
import pytest
from scipy import stats


def _cloud(orc, ncx, ncy, per_cell, ms, mt, rng, vscale=1e5, same=False):
    """A MirSim with one (or two) species holding per_cell particles per cell."""
    nx, ny = ncx + 1, ncy + 1
    sim = orc.MirSim(nx, 1e-3, T=8, sort_every=0, ny=ny, dy=1e-3)
    n = ncx * ncy * per_cell
    for k, m in enumerate((ms,) if same else (ms, mt)):
        ci = np.repeat(np.arange(ncx * ncy), per_cell)
        cell = ((ci % ncx) | ((ci // ncx) << 16)).astype(np.int32)
        sp = orc.MirSpecies(2, n, -1.0 if k == 0 else 0.0, m, 1.0, bnd=(2, 2))
        v = rng.normal(0.0, vscale, (n, 3)).astype(np.float32)
        sp.set_state(cell, rng.random(n).astype(np.float32), rng.random(n).astype(np.float32),
                     v[:, 0], v[:, 1], v[:, 2], np.arange(n, dtype=np.uint32))
        sim.add_species(sp)
    return sim


def _v_by_id(sp):
    """Velocities (FP64) indexed by particle id (the DSMC run re-sorts)."""
    v = np.zeros((sp.np, 3))
    v[sp.id[:sp.np]] = np.stack([sp.vx[:sp.np], sp.vy[:sp.np], sp.vz[:sp.np]], 1)
    return v


def _nc(Na, Nb, W, dx, dy, dt, sgmax, same):
    na = Na * W / (dx * dy)
    nc = na * Nb * dt * sgmax / (1.0 + 1.0)
    return nc if same else 2.0 * nc


def test_trial_count_and_remainder(orc):
    """sigma(g) g >= sgmax for every pair (table flat from a node below every
    relative speed): every trial is accepted, so the per-cell counts are
    floor(Nc + remainder), with the remainder carried to the next call."""
    rng = np.random.default_rng(0)
    ncx, ncy, per = 6, 5, 3
    sim = _cloud(orc, ncx, ncy, per, 9.1e-31, 2.7e-26, rng)
    # sgmax = sigma(node 0) * g(node 0) = 1e-19 * 1.0: relative speeds ~1e5 >> 1
    k = sim.add_dsmc(0, 1, [1.0, 2.0], [1e-19, 1e-19])
    dt = 0.3 / (_nc(per, per, 1.0, 1e-3, 1e-3, 1.0, 1e-19, False))  # Nc = 0.3 per call
    want_rem = np.zeros(ncx * ncy)
    total = 0
    for call in range(7):
        ev = sim.dsmc_run(k, dt)
        nc = _nc(per, per, 1.0, 1e-3, 1e-3, dt, 1e-19, False) + want_rem
        got = np.floor(nc)
        want_rem = nc - got
        assert ev == int(got.sum()), call
        total += ev
        assert np.array_equal(sim.dsmcs[k]["rem"], want_rem), call
    nu = sim.dsmcs[k]["nu"].reshape(ncy + 1, ncx + 1)
    assert nu[:, -1].sum() == 0 and nu[-1, :].sum() == 0  # cells sit at lower-left nodes
    assert nu.sum() == total and total == 2 * ncx * ncy  # 7 x 0.3 = 2.1 per cell


def test_cells_with_fewer_than_two_particles_are_skipped(orc):
    rng = np.random.default_rng(1)
    sim = _cloud(orc, 4, 4, 1, 9.1e-31, 2.7e-26, rng)
    k = sim.add_dsmc(0, 1, [1.0, 2.0], [1e-19, 1e-19])
    assert sim.dsmc_run(k, 1.0) == 0
    assert not np.any(sim.dsmcs[k]["rem"])


@pytest.mark.parametrize("same", [False, True])
def test_collisions_conserve_momentum_and_energy(orc, same):
    rng = np.random.default_rng(2)
    ms, mt = 9.1e-31, 3.2e-30
    ncx, ncy, per = 4, 3, 5
    sim = _cloud(orc, ncx, ncy, per, ms, ms if same else mt, rng, same=same)
    k = sim.add_dsmc(0, 0 if same else 1, [1.0, 2.0], [1e-19, 1e-19])
    sps = sim.species if not same else sim.species[:1]
    masses = [ms] if same else [ms, mt]

    def per_cell():
        p = np.zeros((ncx * ncy, 3))
        e = np.zeros(ncx * ncy)
        for sp, m in zip(sps, masses):
            c = sp.cell[:sp.np]
            key = (c & 0xFFFF) + ncx * (c >> 16)
            v = np.stack([sp.vx[:sp.np], sp.vy[:sp.np], sp.vz[:sp.np]], 1).astype(np.float64)
            np.add.at(p, key, m * v)
            np.add.at(e, key, 0.5 * m * (v * v).sum(1))
        return p, e

    p0, e0 = per_cell()
    v0 = _v_by_id(sps[0])
    dt = 40.0 / _nc(per, per, 1.0, 1e-3, 1e-3, 1.0, 1e-19, same)  # ~40 collisions per cell
    ev = sim.dsmc_run(k, dt)
    assert ev >= 30 * ncx * ncy
    assert np.mean(np.any(_v_by_id(sps[0]) != v0, axis=1)) > 0.9
    p1, e1 = per_cell()
    # each collision rounds six FP32 velocities: relative 2^-24 per component
    pscale = np.abs(p0).max() + 1e-30
    assert np.allclose(p1, p0, rtol=0, atol=2e-5 * pscale)
    assert np.allclose(e1, e0, rtol=2e-5)


def test_post_collision_relative_velocity_is_isotropic(orc):
    """One accepted collision per cell between two pairs: the changed pair's
    relative velocity keeps |g| and points along (B, A cos C, A sin C) with B
    uniform on [-1, 1] and C uniform on [0, 2 pi)."""
    rng = np.random.default_rng(3)
    ncx, ncy = 40, 40
    ms = mt = 1.0
    sim = _cloud(orc, ncx, ncy, 2, ms, mt, rng)
    k = sim.add_dsmc(0, 1, [1.0, 2.0], [1e-19, 1e-19])
    A, B = sim.species
    va0, vb0 = _v_by_id(A), _v_by_id(B)
    dt = 1.5 / _nc(2, 2, 1.0, 1e-3, 1e-3, 1.0, 1e-19, False)  # one trial per cell
    assert sim.dsmc_run(k, dt) == ncx * ncy
    va, vb = _v_by_id(A), _v_by_id(B)  # id 2t, 2t+1 sit in cell t
    cos_b, azim = [], []
    for t in range(ncx * ncy):
        ia = [p for p in (2 * t, 2 * t + 1) if not np.array_equal(va[p], va0[p])]
        ib = [p for p in (2 * t, 2 * t + 1) if not np.array_equal(vb[p], vb0[p])]
        assert len(ia) == 1 and len(ib) == 1, t
        g0 = va0[ia[0]] - vb0[ib[0]]
        g1 = va[ia[0]] - vb[ib[0]]
        assert np.isclose(np.linalg.norm(g1), np.linalg.norm(g0), rtol=1e-6)
        u = g1 / np.linalg.norm(g1)
        cos_b.append(u[0])
        azim.append(np.arctan2(u[2], u[1]) % (2 * np.pi))
    assert stats.kstest(cos_b, stats.uniform(loc=-1, scale=2).cdf).pvalue > 1e-3
    assert stats.kstest(azim, stats.uniform(loc=0, scale=2 * np.pi).cdf).pvalue > 1e-3


def test_acceptance_rule_rate(orc):
    """sigma flat at sigma0 with its first node at g_ref: P = g / g_ref below
    g_ref.  The accepted fraction of the trials matches E[min(1, g/g_ref)]
    over the cells' pairs (z-test)."""
    rng = np.random.default_rng(4)
    ncx, ncy, per = 30, 30, 4
    sim = _cloud(orc, ncx, ncy, per, 1.0, 1.0, rng, vscale=1.0)
    g_ref = 3.0
    k = sim.add_dsmc(0, 1, [g_ref, 10.0], [1e-19, 1e-19])
    A, B = sim.species
    va = np.stack([A.vx[:A.np], A.vy[:A.np], A.vz[:A.np]], 1).astype(np.float64)
    vb = np.stack([B.vx[:B.np], B.vy[:B.np], B.vz[:B.np]], 1).astype(np.float64)
    p = []
    for t in range(ncx * ncy):
        ga = va[t * per:(t + 1) * per][:, None, :] - vb[t * per:(t + 1) * per][None, :, :]
        g = np.linalg.norm(ga, axis=2)
        # equal device indices are redrawn, and both species hold the cell at
        # the same positions: the pairs are the off-diagonal ones
        off = ~np.eye(per, dtype=bool)
        p.append(np.minimum(1.0, g[off] / g_ref).mean())
    ntr = 1
    dt = (ntr + 0.5) / _nc(per, per, 1.0, 1e-3, 1e-3, 1.0, 1e-19 * g_ref, False)
    ev = sim.dsmc_run(k, dt)
    mean_p = np.mean(p)
    n = ncx * ncy * ntr
    z = (ev - n * mean_p) / np.sqrt(n * mean_p * (1 - mean_p))
    assert abs(z) < 4.0, (ev, n * mean_p, z)


def test_index_redraw_only_between_species(orc):
    """With one species as source and target the same particle may be drawn
    twice (g = 0, never accepted), so about half the trials of two-particle
    cells are rejected; with two species equal indices are redrawn (the
    acceptance-rate test depends on it)."""
    rng = np.random.default_rng(5)
    sim_same = _cloud(orc, 3, 3, 2, 1.0, 1.0, rng, same=True)
    k = sim_same.add_dsmc(0, 0, [1.0, 2.0], [1e-19, 1e-19])
    dt = 50.0 / _nc(2, 2, 1.0, 1e-3, 1e-3, 1.0, 1e-19, True)
    ev = sim_same.dsmc_run(k, dt)
    # 50 trials per cell over 2 x 2 equally likely (source, target) pairs:
    # about half pick the same particle and are rejected
    assert 0.3 * 50 * 9 < ev < 0.7 * 50 * 9
