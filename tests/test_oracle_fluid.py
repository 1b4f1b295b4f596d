"""The oracle's restatement of calculate_advection_diffusion
(FiniteDifferenceMethod/src/advection_diffusion.jl:4-38) pinned by closed
forms of the discrete operators (each chosen so that a plausible slip, such
as dividing the y terms by dy, a dx^2 for the reference's dx dy, or a sign,
fails it) and by a literal loop transcription on random data."""
import numpy as np

import oracle as O


def _loops(n, D, v, dh, dt):
    """the reference's loops, one node at a time"""
    d = np.zeros_like(n)
    if n.ndim == 1:
        dx = dh[0]
        for i in range(1, n.shape[0] - 1):
            d[i] += dt * D * (n[i - 1] - 2 * n[i] + n[i + 1]) / dx ** 2
        return d
    dx, dy = dh
    vx, vy = v[..., 0], v[..., 1]
    for i in range(1, n.shape[0] - 1):
        for j in range(1, n.shape[1] - 1):
            d[i, j] += dt * D * (n[i - 1, j] + n[i, j - 1] - 4 * n[i, j] + n[i + 1, j]
                                 + n[i, j + 1]) / (dx * dy)
            d[i, j] -= dt * n[i, j] * (vx[i + 1, j] - vx[i - 1, j] + vy[i, j + 1]
                                       - vy[i, j - 1]) / (2 * dx)
            d[i, j] -= dt * (vx[i, j] * (n[i + 1, j] - n[i - 1, j])
                             + vy[i, j] * (n[i, j + 1] - n[i, j - 1])) / dx
    return d


def test_one_d_sine_mode():
    nx, dx, dt, D, k = 33, 0.25, 1e-3, 0.7, 3
    i = np.arange(nx)
    n = np.sin(np.pi * k * i / (nx - 1))
    d = O.ref_advection_diffusion(n, D, np.zeros((nx, 3)), (dx,), dt)
    lam = 2.0 * np.cos(np.pi * k / (nx - 1)) - 2.0
    assert d[0] == 0.0 and d[-1] == 0.0
    assert np.allclose(d[1:-1], dt * D * lam / dx ** 2 * n[1:-1], rtol=1e-12, atol=1e-15)


def test_two_d_uniform_density_linear_drift():
    # n = c, v = (a x, b y): only the divergence term acts, divided by 2 dx
    # for both axes as the reference writes it (so dx != dy tells them apart)
    nx, ny, dx, dy, dt, c, a, b = 9, 7, 0.5, 0.2, 1e-2, 3.0, 0.4, -1.3
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    v = np.stack([a * X, b * Y, np.zeros_like(X)], axis=-1)
    d = O.ref_advection_diffusion(np.full((nx, ny), c), 0.9, v, (dx, dy), dt)
    want = -dt * c * (2 * a * dx + 2 * b * dy) / (2 * dx)
    assert np.allclose(d[1:-1, 1:-1], want, rtol=1e-12)
    assert not d[0].any() and not d[-1].any() and not d[:, 0].any() and not d[:, -1].any()


def test_two_d_linear_density_constant_drift():
    # n = alpha x + beta y, v = (V, W): only the upwind-free gradient term acts
    nx, ny, dx, dy, dt, al, be, V, W = 8, 6, 0.3, 0.7, 1e-2, 2.0, -0.5, 1.5, 0.25
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    v = np.stack([np.full_like(X, V), np.full_like(X, W), np.zeros_like(X)], axis=-1)
    d = O.ref_advection_diffusion(al * X + be * Y, 0.0, v, (dx, dy), dt)
    want = -dt * (V * 2 * al * dx + W * 2 * be * dy) / dx
    assert np.allclose(d[1:-1, 1:-1], want, rtol=1e-12)


def test_two_d_quadratic_density_diffusion():
    # n = x^2 + y^2: the five-point sum is 2 dx^2 + 2 dy^2, divided by dx dy
    nx, ny, dx, dy, dt, D = 7, 9, 0.5, 0.25, 1e-3, 0.6
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dy, indexing="ij")
    d = O.ref_advection_diffusion(X ** 2 + Y ** 2, D, np.zeros((nx, ny, 3)), (dx, dy), dt)
    want = dt * D * (2 * dx ** 2 + 2 * dy ** 2) / (dx * dy)
    assert np.allclose(d[1:-1, 1:-1], want, rtol=1e-12)


def test_restatement_equals_literal_loops():
    rng = np.random.default_rng(4)
    for shape, dh in (((13,), (0.3,)), ((9, 11), (0.4, 0.15))):
        n = rng.random(shape)
        v = rng.normal(size=shape + (3,))
        a = O.ref_advection_diffusion(n, 0.35, v, dh, 2e-3)
        b = _loops(n, 0.35, v, dh, 2e-3)
        assert np.array_equal(a, b), shape


def test_refsim_fluid_moves_with_the_previous_field():
    # a charged 2-D fluid in a box with phi = 0 walls: the first step sees
    # E = 0 (pure diffusion), later steps the field of the previous rho
    nx, dx = 9, 0.5
    sim = O.RefSim(2, nx, dx, ny=nx, dy=dx, eps0=1.0)
    tags = np.zeros((nx, nx), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    sim.set_tags(tags.T.ravel())
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(nx) * dx, indexing="ij")
    n0 = np.exp(-((X - 2.0) ** 2 + (Y - 2.0) ** 2))
    f = sim.add_fluid(1.0, 2.0, 0.1, n0.T.ravel())
    sim.step(0.01)
    d1 = O.ref_advection_diffusion(n0, 0.1, np.zeros((nx, nx, 3)), (dx, dx), 0.01)
    assert np.array_equal(sim.fluids[f]["n"], (n0 + d1).T.ravel())
    E = sim.E.reshape(3, -1)
    v = np.stack([(E[k].reshape(nx, nx).T * 0.01) * 0.5 for k in range(3)], axis=-1)
    n1 = n0 + d1
    sim.step(0.01)
    d2 = O.ref_advection_diffusion(n1, 0.1, v, (dx, dx), 0.01)
    assert np.array_equal(sim.fluids[f]["n"], (n1 + d2).T.ravel())
    assert np.allclose(sim.rho, sim.fluids[f]["n"], rtol=0, atol=0)
