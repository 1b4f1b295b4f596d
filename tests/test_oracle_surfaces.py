"""The surface tracker, floating-electrode rows and RLC circuit restated in
oracle/surfaces.py, pinned by closed forms; then the device's tracer twin
(mir_push_deposit_track) against the restated tracker."""
import math

import numpy as np
import pytest

from oracle import surfaces as S


def _free_push(x, v, dt):
    """push_in_cartesian! with E = B = 0: x += v dt."""
    def push(xx, vv):
        return xx + vv[:, :2] * dt, vv
    return push


def test_absorbing_boundary_and_reach():
    nx, ny, dh, dt = 6, 5, 1.0, 1.0
    st = S.SurfaceTracker(nx, ny, dh)
    x = np.array([[0.3, 2.5], [0.3, 2.5], [4.8, 1.5]])
    v = np.array([[-0.5, 0.0, 0.0], [-0.2, 0.0, 0.0], [0.0, -0.4, 0.0]])
    gone = st.step(x, v, 3, dt, 1.0, 1.0, _free_push(x, v, dt))
    # particle 0 crosses x = 0, particle 1 stops at 0.1, particle 2 stays above y = 0
    assert gone == [0]
    assert np.allclose(x[1], [0.1, 2.5]) and np.allclose(x[2], [4.8, 1.1])


def test_reflective_wall_mirrors_the_rest_of_the_step():
    nx, ny, dh, dt = 8, 6, 0.5, 1.0
    st = S.SurfaceTracker(nx, ny, dh)
    wall = np.zeros((nx, ny), bool)
    wall[4, :] = True  # x = 2.0
    st.track_surface(wall, S.REFLECT)
    x = np.array([[1.8, 1.3]])
    v = np.array([[0.7, 0.1, 3.0]])
    gone = st.step(x, v, 1, dt, 1.0, 1.0, _free_push(x, v, dt))
    assert gone == []
    # unreflected end 2.5 -> mirrored about 2.0 -> 1.5; y and v_y, v_z unchanged
    assert np.allclose(x[0], [1.5, 1.4]) and np.allclose(v[0], [-0.7, 0.1, 3.0])


def test_reflective_corner_reflects_twice():
    nx, ny, dh, dt = 8, 8, 1.0, 1.0
    st = S.SurfaceTracker(nx, ny, dh)
    box = np.zeros((nx, ny), bool)
    box[5, :] = True  # x = 5
    box[:, 5] = True  # y = 5
    st.track_surface(box, S.REFLECT)
    x = np.array([[4.6, 4.7]])
    v = np.array([[0.9, 0.8, 0.0]])
    st.step(x, v, 1, dt, 1.0, 1.0, _free_push(x, v, dt))
    assert np.allclose(x[0], [10.0 - 5.5, 10.0 - 5.5]) and np.allclose(v[0, :2], [-0.9, -0.8])


def test_untracked_cells_cross_surfaces_unseen():
    """track! only follows particles that start next to a surface: one that
    starts two cells away and jumps over a reflective line is not reflected
    (the reference's 'too fast' case)."""
    nx, ny, dh, dt = 10, 6, 1.0, 1.0
    st = S.SurfaceTracker(nx, ny, dh)
    wall = np.zeros((nx, ny), bool)
    wall[5, :] = True
    st.track_surface(wall, S.REFLECT)
    x = np.array([[2.5, 2.5]])
    v = np.array([[3.5, 0.0, 0.0]])
    st.step(x, v, 1, dt, 1.0, 1.0, _free_push(x, v, dt))
    assert np.allclose(x[0], [6.0, 2.5]) and v[0, 0] == 3.5


def test_floating_electrode_collects_charge():
    nx, ny, dh, dt = 6, 6, 1.0, 1.0
    st = S.SurfaceTracker(nx, ny, dh)
    mask = np.zeros((nx, ny), bool)
    mask[0, 1:5] = True
    e = S.Electrode(False, S.electrode_area(mask, dh, dh), sigma=0.25)
    assert e.area == 3.0
    st.track_surface(mask, e)
    x = np.array([[0.4, 2.5], [0.4, 3.5], [0.4, 0.5]])
    v = np.array([[-1.0, 0.0, 0.0]] * 3)
    gone = st.step(x, v, 3, dt, -2.0, 3.0, _free_push(x, v, dt))
    # particles 0 and 1 hit the electrode, particle 2 the default absorbing wall
    assert gone == [2, 1, 0]
    assert e.dq == 2 * (-2.0 * 3.0) and e.sigma == 0.25 + 2 * (-6.0 / 3.0)


def test_twin_tracer_against_restated_tracker(orc):
    """The device's arithmetic (cell + FP32 offset, FP64 trace) against the
    restated reference on 20,000 particles in the 07_boundaries geometry:
    same absorbed set, same electrode charge, positions and velocities within
    FP32 rounding."""
    nx, ny, dh = 11, 11, 0.1
    bcs = S.boundaries_geometry(nx, ny)
    st = S.SurfaceTracker(nx, ny, dh)
    e1 = S.Electrode(False, S.electrode_area(bcs == 1, dh, dh))
    e2 = S.Electrode(False, S.electrode_area(bcs == 2, dh, dh))
    e3 = S.Electrode(True, S.electrode_area(bcs == 3, dh, dh))
    st.track_surface(bcs == 1, e1)
    st.track_surface(bcs == 2, e2)
    st.track_surface(bcs == 3, e3)
    st.track_surface(bcs == 4, S.REFLECT)
    codes, tracked = S.tracker_table(st, nx, ny, [e1, e2])
    rng = np.random.default_rng(0)
    n, dt = 20000, 1e-8
    ci = rng.integers(0, nx - 1, n)
    cj = rng.integers(0, ny - 1, n)
    ox = rng.random(n).astype(np.float32)
    oy = rng.random(n).astype(np.float32)
    vel = rng.normal(0.0, 0.8 * dh / dt, (n, 3)).astype(np.float32)
    sp = orc.MirSpecies(2, n, -1.6e-19, 9.1e-31, 5.0, bnd=(2, 2))
    sp.set_state((ci | (cj << 16)).astype(np.int32), ox, oy, vel[:, 0], vel[:, 1], vel[:, 2],
                 np.arange(n, dtype=np.uint32))
    sim = orc.MirSim(nx, dh, T=8, sort_every=0, ny=ny, dy=dh)
    sim.add_species(sp)
    sim.set_tracker(codes, tracked)
    sim.set_electrodes([0.0, 0.0], [e1.area, e2.area])
    sim.push(0, dt)
    sim.electrode_update(dt, False)
    x = np.column_stack([(ci + ox.astype(np.float64)) * dh, (cj + oy.astype(np.float64)) * dh])
    v = vel.astype(np.float64)
    gone = st.step(x, v, n, dt, -1.6e-19, 5.0, _free_push(x, v, dt))
    # the reference removes absorbed particles and then applies after_push
    # (here discard!); the twin marks both the same way (cell -1)
    out = (x[:, 0] < 0) | (x[:, 0] >= (nx - 1) * dh) | (x[:, 1] < 0) | (x[:, 1] >= (ny - 1) * dh)
    ref_dead = np.zeros(n, bool)
    ref_dead[gone] = True
    ref_dead |= out
    dead = sp.cell[:n] < 0
    # a particle ending within FP32 rounding of a cell edge may land on
    # either side: allow a handful
    assert np.sum(dead != ref_dead) <= 5, np.sum(dead != ref_dead)
    ok = ~dead & ~ref_dead
    xi = (sp.cell[:n] & 0xFFFF) + sp.ox[:n].astype(np.float64)
    yj = (sp.cell[:n] >> 16) + sp.oy[:n].astype(np.float64)
    assert np.allclose(xi[ok] * dh, x[ok, 0], atol=2e-6 * nx * dh)
    assert np.allclose(yj[ok] * dh, x[ok, 1], atol=2e-6 * ny * dh)
    assert np.array_equal(sp.vx[:n][ok], v[ok, 0].astype(np.float32))
    assert np.array_equal(sp.vy[:n][ok], v[ok, 1].astype(np.float32))
    # electrode charges: the same particles, summed per species in the twin
    assert math.isclose(sim.el_dq[0], e1.dq, rel_tol=1e-12) and e1.dq != 0.0
    assert math.isclose(sim.el_dq[1], e2.dq, rel_tol=1e-12) and e2.dq != 0.0
    assert np.any(v[:, 0] != vel[:, 0].astype(np.float64))  # reflections happened


def test_floating_wall_electrode_gives_the_linear_potential():
    """apply_neumann rows (generalized_poisson.jl:237-269) of a floating left
    wall over a grounded right wall, no charge: phi = sigma (L - x), the
    potential whose slope the rows prescribe.  With the wall's two corner
    nodes fixed at that value the segment-end rows (c != d) hold it too."""
    nx, ny, dh, sigma = 9, 7, 0.25, 3.0
    L = (nx - 1) * dh
    xg = np.arange(nx) * dh
    want = np.tile(sigma * (L - xg), ny)
    left = np.zeros((nx, ny), bool)
    left[0, :] = True
    right = np.zeros((nx, ny), bool)
    right[-1, :] = True
    A, b, rho, _ = S.electrode_system(nx, ny, dh, dh, fixed=[(right, 0.0)],
                                      floating=[(left, sigma)])
    phi = S.electrode_potential(A, b, rho, np.zeros(nx * ny), 1.0)
    assert np.allclose(phi, want, rtol=0, atol=1e-12 * sigma * L)
    seg = left.copy()
    seg[0, 0] = seg[0, -1] = False
    corners = np.zeros((nx, ny), bool)
    corners[0, 0] = corners[0, -1] = True
    A, b, rho, _ = S.electrode_system(nx, ny, dh, dh,
                                      fixed=[(right, 0.0), (corners, sigma * L)],
                                      floating=[(seg, sigma)])
    phi = S.electrode_potential(A, b, rho, np.zeros(nx * ny), 1.0)
    assert np.allclose(phi, want, rtol=0, atol=1e-12 * sigma * L)
    # a planted error in the segment-end coefficient breaks it
    A2 = A.copy()
    k = 0 + nx * 1  # node (1, 2): c != d
    A2[k, nx * ny] *= 1.01
    bb = b.copy()
    bb[:nx * ny][rho] = 0.0
    assert not np.allclose(np.linalg.solve(A2, bb)[:nx * ny], want, atol=1e-6)


def test_rl_step_response():
    """advance_circuit! with C = 0 and V = V0: L di/dt + R i = -V0, so
    i(t) = -(V0/R)(1 - exp(-R t / L)); the scheme converges to it at first
    order or better as dt shrinks."""
    R, L, V0 = 50.0, 1e-6, 2.0

    def err(dt):
        c = S.CircuitRLC(R, L, 0.0, v0=V0)
        for _ in range(int(round(5 * L / R / dt))):
            c.advance(0.0, dt)
        return abs(c.i + (V0 / R) * (1 - math.exp(-R * c.t / L)))

    e1, e2 = err(1e-9), err(1e-10)
    assert e2 < 2e-3 * V0 / R and e2 < e1 / 5, (e1, e2)


def test_rc_discharge():
    """C > 0, L small: the capacitor charge follows the RLC solution; with
    R^2 > 4 L / C it decays without oscillation towards -V0 C."""
    R, L, Cc, V0, dt = 100.0, 1e-7, 1e-9, 1.0, 1e-10
    c = S.CircuitRLC(R, L, Cc, v0=V0)
    for _ in range(int(20 * R * Cc / dt)):
        c.advance(0.0, dt)
    assert abs(c.q - (-V0 * Cc)) <= 1e-3 * V0 * Cc
