"""The axial (r, z) restatements of the oracle against closed forms: the node
volumes of cell_volume(::AxialGrid{2}) (RegularGrids.jl:40-53), the
axisymmetric Poisson matrix (generalized_poisson.jl:70-199) on manufactured
potentials, and transform_from_cartesian_to_cylindrical!
(ParticleInCell/src/pic/pushers.jl:52-66); then the FP32 twin of the RZ push
against the FP64 reference operators."""
import numpy as np
import pytest


@pytest.mark.parametrize("nr,nz,dr,dz", [(5, 4, 0.1, 0.25), (17, 9, 1e-3, 2e-3), (2, 2, 1.0, 1.0)])
def test_axial_cell_volume_closed_forms(orc, nr, nz, dr, dz):
    V = orc.cell_volume_axial(nr, nz, dr, dz).reshape(nz, nr)
    i = np.arange(nr)
    want = 2.0 * np.pi * i * dr * dr * dz          # annulus of radius i dr, width dr
    want[0] = np.pi * (0.5 * dr) ** 2 * dz         # disc on the axis
    R = (nr - 1) * dr
    want[nr - 1] = np.pi * (R ** 2 - (R - 0.5 * dr) ** 2) * dz  # half annulus at the wall
    for j in range(nz):
        w = want * (0.5 if j in (0, nz - 1) else 1.0)
        assert np.allclose(V[j], w, rtol=1e-13, atol=0), j
    # the node volumes tile the cylinder of radius R and length (nz-1) dz
    assert np.isclose(V.sum(), np.pi * R ** 2 * (nz - 1) * dz, rtol=1e-13)
    # a periodic z edge is not halved
    Vp = orc.cell_volume_axial(nr, nz, dr, dz, (1, 0)).reshape(nz, nr)
    assert np.array_equal(Vp[0], V[0] * 2.0) and np.array_equal(Vp[-1], V[-1])


def _rz_matrix(orc, nr, nz, dr, dz):
    A, _, _ = orc.poisson_matrix_rz(nr, nz, dr, dz)
    return A.reshape(nr * nz, nr * nz, order="F")


@pytest.mark.parametrize("nr,nz,dr,dz", [(9, 7, 0.5, 0.25), (6, 11, 2e-3, 1e-3)])
def test_axial_operator_on_manufactured_potentials(orc, nr, nz, dr, dz):
    """phi = r^2 has Laplacian 4 and phi = z^2 has 2: the reference's rows are
    exact there on interior nodes; the axis, wall and z-edge rows are the
    mirrored one-sided forms, checked against their hand values."""
    A = _rz_matrix(orc, nr, nz, dr, dz)
    k = np.arange(nr * nz)
    i, j = k % nr, k // nr
    r, z = i * dr, j * dz
    Ar2 = A @ (r ** 2)
    Az2 = A @ (z ** 2)
    # radial part on r^2: interior rows 4; axis row (r1^2 - r0^2)/dr^2 = 1;
    # wall row (r_{n-2}^2 - r_{n-1}^2)/dr^2 = -(2 nr - 3)
    want_r = np.where(i == 0, 1.0, np.where(i == nr - 1, -(2.0 * nr - 3.0), 4.0))
    assert np.allclose(Ar2, want_r, rtol=0, atol=1e-9 * (2 * nr)), np.max(np.abs(Ar2 - want_r))
    want_z = np.where(j == 0, 1.0, np.where(j == nz - 1, -(2.0 * nz - 3.0), 2.0))
    assert np.allclose(Az2, want_z, rtol=0, atol=1e-9 * (2 * nz)), np.max(np.abs(Az2 - want_z))
    # constants are in the null space (every row sums to zero: natural edges)
    assert np.allclose(A @ np.ones(nr * nz), 0.0, atol=1e-9 / min(dr, dz) ** 2)
    # a planted sign error in the derivative term breaks the r^2 check
    bad = A.copy()
    m = (i > 0) & (i < nr - 1)
    for kk in k[m]:
        bad[kk, kk + 1] -= 1.0 / dr / r[kk]
        bad[kk, kk - 1] += 1.0 / dr / r[kk]
    assert not np.allclose(bad @ (r ** 2), want_r, atol=1e-6)


def test_axial_operator_weighted_symmetry_and_sparse_twin(orc):
    """Row weights w_0 = 1/2, w_i = i, w_{nr-1} = nr - 3/2 make W A symmetric
    (the form the device's CG solves); the sparse assembly equals the dense
    one, Dirichlet rows included."""
    nr, nz, dr, dz = 8, 6, 0.3, 0.2
    A = _rz_matrix(orc, nr, nz, dr, dz)
    k = np.arange(nr * nz)
    i = k % nr
    w = np.where(i == 0, 0.5, np.where(i == nr - 1, nr - 1.5, i.astype(float)))
    WA = w[:, None] * A
    assert np.allclose(WA, WA.T, rtol=0, atol=1e-12 * np.abs(WA).max())
    # negative semi-definite with the constant null vector only
    ev = np.linalg.eigvalsh(0.5 * (WA + WA.T))
    assert ev.max() < 1e-9 * np.abs(ev).max() and np.sum(np.abs(ev) < 1e-9 * np.abs(ev).max()) == 1
    tags = np.zeros((nz, nr), np.int8)
    tags[:, -1] = 1
    tags[0, :] = 2
    tv = [0.0, 1.5, -0.5]
    S, dmask, dval = orc.poisson_sparse_rz(nr, nz, dr, dz, tags.ravel(), tv)
    Ad, b, isdof = orc.poisson_matrix_rz(nr, nz, dr, dz, dirichlet=[
        ((tags == 1).T, 1.5), ((tags == 2).T, -0.5)])
    Ad = Ad.reshape(nr * nz, nr * nz, order="F")
    assert np.allclose(S.toarray(), Ad, rtol=1e-14, atol=0)
    assert np.array_equal(dmask, isdof == 0)
    assert np.array_equal(dval[dmask], b[dmask])


def test_rz_transform_closed_forms(orc):
    rng = np.random.default_rng(7)
    n, dt = 1000, 1e-9
    x = np.column_stack([rng.uniform(1e-3, 5e-2, n), rng.uniform(0, 1e-2, n)])
    v = rng.normal(0, 1e6, (n, 3))
    x0, v0 = x.copy(), v.copy()
    orc.rz_transform(x, v, dt)
    assert np.allclose(x[:, 0], np.hypot(x0[:, 0], dt * v0[:, 2]), rtol=1e-15)
    assert np.array_equal(x[:, 1], x0[:, 1]) and np.array_equal(v[:, 1], v0[:, 1])
    assert np.allclose(v[:, 0] ** 2 + v[:, 2] ** 2, v0[:, 0] ** 2 + v0[:, 2] ** 2, rtol=1e-13)
    # the rotation angle is the one the particle swept: tan = dt v_theta / r
    sn = dt * v0[:, 2] / x[:, 0]
    cs = x0[:, 0] / x[:, 0]
    assert np.allclose(v[:, 0], cs * v0[:, 0] + sn * v0[:, 2], rtol=1e-12, atol=1e-6)
    # v_theta = 0: nothing turns
    x1, v1 = x0.copy(), v0.copy()
    v1[:, 2] = 0.0
    x2, v2 = x1.copy(), v1.copy()
    orc.rz_transform(x2, v2, dt)
    assert np.array_equal(x2, x1) and np.array_equal(v2, v1)
    # on the axis (r = 0 after the step, v_theta = 0) sin is 0, not NaN
    xa = np.array([[0.0, 0.5]])
    va = np.array([[3.0, 1.0, 0.0]])
    orc.rz_transform(xa, va, dt)
    assert np.array_equal(xa, [[0.0, 0.5]]) and np.array_equal(va, [[3.0, 1.0, 0.0]])
    # crossing the axis: the reference takes cos = +sqrt(1 - sin^2), so v_r
    # keeps its (inward) sign where the true rotation would flip it
    xc = np.array([[-2e-3, 0.0]])
    vc = np.array([[-1e6, 0.0, 1e5]])
    orc.rz_transform(xc, vc, dt)
    assert xc[0, 0] > 0 and vc[0, 0] < 0


def test_rz_twin_push_against_reference_step(orc):
    """mir_push_deposit_rz (the device's arithmetic) against ref_push followed
    by ref_rz_transform on the same particles and E, in FP32 tolerance."""
    nr, nz, dr, dz = 33, 17, 1e-3, 1e-3
    rng = np.random.default_rng(3)
    n, dt, qm = 4000, 2e-11, -1.758820e11
    cell_i = rng.integers(0, nr - 1, n)
    cell_j = rng.integers(0, nz - 1, n)
    ox = rng.random(n).astype(np.float32)
    oy = rng.random(n).astype(np.float32)
    vel = rng.normal(0, 3e5, (n, 3)).astype(np.float32)
    E = rng.normal(0, 1e4, (nz * nr, 2)).astype(np.float32)
    sp = orc.MirSpecies(2, n, -1.602e-19, 9.109e-31, 1.0, bnd=(2, 2))
    sp.set_state((cell_i | (cell_j << 16)).astype(np.int32), ox, oy, vel[:, 0], vel[:, 1],
                 vel[:, 2], np.arange(n, dtype=np.uint32))
    sim = orc.MirSim(nr, dr, T=8, sort_every=0, ny=nz, dy=dz, axial=True)
    sim.add_species(sp)
    sim.E[:] = E.ravel()
    # reference: absolute FP64 positions, gathered E, Cartesian push, transform
    x = np.column_stack([(cell_i + ox.astype(np.float64)) * dr, (cell_j + oy.astype(np.float64)) * dz])
    v = vel.astype(np.float64)
    Eg = np.zeros(nr * nz * 3)
    Eg[:nr * nz] = E[:, 0]
    Eg[nr * nz:2 * nr * nz] = E[:, 1]
    xf = np.ascontiguousarray(x.T.ravel())
    vf = np.ascontiguousarray(v.T.ravel())
    pE = orc.grid_to_particle(2, xf, nr, nz, dr, dz, Eg)
    orc.ref_push(xf, vf, pE, np.zeros_like(pE), qm, dt, 2)
    xr, vr = xf.reshape(2, n).T.copy(), vf.reshape(3, n).T.copy()
    orc.rz_transform(xr, vr, dt)
    sp.q, sp.m = qm, 1.0
    sim.push(0, dt)
    keep = (xr[:, 0] < (nr - 1) * dr) & (xr[:, 1] >= 0) & (xr[:, 1] < (nz - 1) * dz)
    alive = sp.cell[:n] >= 0
    assert np.array_equal(alive, keep)
    ci, cj = sp.cell[:n] & 0xFFFF, sp.cell[:n] >> 16
    rpos = (ci + sp.ox[:n].astype(np.float64)) * dr
    zpos = (cj + sp.oy[:n].astype(np.float64)) * dz
    assert np.allclose(rpos[alive], xr[alive, 0], rtol=0, atol=1e-6 * dr * nr)
    assert np.allclose(zpos[alive], xr[alive, 1], rtol=0, atol=1e-6 * dz * nz)
    for c, a in enumerate((sp.vx, sp.vy, sp.vz)):
        assert np.allclose(a[:n][alive], vr[alive, c], rtol=2e-6, atol=1e-6 * np.abs(vr).max())


def test_axial_sparse_solve_equals_dense_lu(orc):
    """potential_rz (free rows, Dirichlet columns moved to the right-hand
    side, sparse LU) is the solution of the reference's full dense system."""
    nr, nz, dr, dz = 10, 7, 2e-3, 1e-3
    tags = np.zeros((nz, nr), np.int8)
    tags[:, -1] = 1
    tags[0, :] = 2
    rng = np.random.default_rng(1)
    rho = rng.normal(0.0, 1e-6, nr * nz)
    eps0 = 8.854e-12
    x = orc.potential_rz(nr, nz, dr, dz, tags, [0.0, 3.0, -7.0], rho, eps0)
    A, b, isdof = orc.poisson_matrix_rz(nr, nz, dr, dz,
                                        dirichlet=[((tags == 1).T, 3.0), ((tags == 2).T, -7.0)])
    ref, sing = orc.potential(A, b, isdof, rho, eps0)
    assert not sing
    assert np.max(np.abs(x - ref)) <= 1e-11 * np.max(np.abs(ref))
    assert np.array_equal(x[tags.ravel() == 1], np.full((tags == 1).sum(), 3.0))
