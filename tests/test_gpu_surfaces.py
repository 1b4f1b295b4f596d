"""The surface tracker, electrodes and RLC circuit on the device
(ParticleInCell/src/pic/surfaces/*.jl, pic/circuit_coupling.jl,
generalized_poisson.jl:217-269, Circuit/src/Circuit.jl:117-136, in the
geometry of problem/07_boundaries.jl) against the CPU twin bit for bit and
the restated reference system."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS0, QE, ME = 8.8541878128e-12, 1.602176634e-19, 9.1093837015e-31


def _geometry_ctx(orc, nx=11, ny=11, dh=0.1, circuit=False):
    """07_boundaries.jl's surfaces on a context: returns ctx, the masks, the
    oracle tracker mirror and the electrode ids."""
    from iskra_amd import Context, _lib as L
    from oracle import surfaces as S
    bcs = S.boundaries_geometry(nx, ny)
    ctx = Context()
    ctx.grid(2, (nx, ny), (dh, dh))
    ctx.tuning(tile=8, chunk=1024, margin=2, sort_every=4)
    ctx.poisson(L.POISSON_PCG, EPS0, None, rtol=1e-13, max_iter=20000)
    flat = lambda m: np.ascontiguousarray(m.T).ravel()  # noqa: E731 (x fastest)
    ids = {}
    ids["driven"] = ctx.electrode(flat(bcs == 1), fixed=False, sigma=1e-9)
    ids["floating"] = ctx.electrode(flat(bcs == 2), fixed=False, sigma=0.0)
    ids["grounded"] = ctx.electrode(flat(bcs == 3), fixed=True, phi=0.0)
    ctx.tracker_surface(flat(bcs == 4), L.SURF_REFLECTIVE)
    st = S.SurfaceTracker(nx, ny, dh)
    e1 = S.Electrode(False, S.electrode_area(bcs == 1, dh, dh), 1e-9)
    e2 = S.Electrode(False, S.electrode_area(bcs == 2, dh, dh), 0.0)
    e3 = S.Electrode(True, S.electrode_area(bcs == 3, dh, dh))
    st.track_surface(bcs == 1, e1)
    st.track_surface(bcs == 2, e2)
    st.track_surface(bcs == 3, e3)
    st.track_surface(bcs == 4, S.REFLECT)
    return ctx, bcs, st, (e1, e2), ids


def test_tracker_table_is_the_reference_lookup(gpu_ready, orc):
    from oracle import surfaces as S
    ctx, bcs, st, (e1, e2), ids = _geometry_ctx(orc)
    try:
        codes, tracked = ctx.tracker_table()
        want_c, want_t = S.tracker_table(st, 11, 11, [e1, e2])
        assert np.array_equal(codes, want_c)
        assert np.array_equal(tracked, want_t)
        for k, e in ((ids["driven"], e1), (ids["floating"], e2)):
            assert ctx.electrode_state(k)["area"] == e.area
    finally:
        ctx.close()


def test_electrode_potential_against_reference_system(gpu_ready, orc):
    """BiCGStab on the rows apply_neumann writes, against the dense solve of
    the reference's system (add_new_dof + apply_neumann + apply_dirichlet)."""
    from iskra_amd import _lib as L
    from oracle import surfaces as S
    ctx, bcs, st, _, ids = _geometry_ctx(orc)
    try:
        rng = np.random.default_rng(3)
        rho = rng.normal(0.0, 1e-9, 121)
        ctx.set_field(L.F_RHO, rho)
        it, rr = ctx.solve(L.WARM_ZERO)
        assert rr <= 1e-13, (it, rr)
        phi = ctx.field(L.F_PHI)
        A, b, rdof, _ = S.electrode_system(11, 11, 0.1, 0.1, fixed=[(bcs == 3, 0.0)],
                                           floating=[(bcs == 1, 1e-9), (bcs == 2, 0.0)])
        ref = S.electrode_potential(A, b, rdof, rho, EPS0)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(phi - ref)) <= 1e-9 * scale, np.max(np.abs(phi - ref)) / scale
        st_d = ctx.electrode_state(ids["driven"])
        assert st_d["phi"] == phi[np.flatnonzero(np.ascontiguousarray((bcs == 1).T).ravel())[0]]
    finally:
        ctx.close()


def test_surfaces_in_time_loop_bitexact(gpu_ready, orc):
    """A thermalised electron beam in 07_boundaries.jl's box with a circuit
    driving the left electrode against the grounded nodes: 24 steps of
    iskra_step against the twin.  Each step: rho bit for bit, the device's
    BiCGStab potential within 1e-8 of the reference system's dense solve and
    handed to the twin, E bit for bit, electrode charges and circuit state bit
    for bit; particle states every 4 steps."""
    from iskra_amd import _lib as L
    from oracle import surfaces as S
    nx = ny = 11
    dh, dt = 0.1, 1e-8
    ctx, bcs, st, (e1, e2), ids = _geometry_ctx(orc)
    try:
        ctx.circuit(R=1.0, L=1e-6, C_=1e-6, v_offset=0.0, v_amplitude=1.0, v_freq=5e6,
                    positive=ids["driven"], negative=ids["grounded"])
        w0 = 5e4
        n = 4000
        s = ctx.species(-QE, ME, w0, n, track_ids=True)
        ctx.boundary(s, [L.BND_WRAP, L.BND_WRAP])
        ctx.sample_maxwellian(s, n, [0.5, 0.5], [3e6] * 3, dx=[0.25, 0.25], seed=3, stream=0)
        codes, tracked = ctx.tracker_table()
        stt = ctx.get_state(s)
        sp = orc.MirSpecies(2, n, -QE, ME, w0, bnd=(1, 1))
        sp.set_state(stt["cell"], stt["ox"], stt["oy"], stt["vx"], stt["vy"], stt["vz"], stt["id"])
        sim = orc.MirSim(nx, dh, T=8, sort_every=4, ny=ny, dy=dh)
        sim.add_species(sp)
        sim.set_tracker(codes, tracked)
        ref_node = int(np.flatnonzero(np.ascontiguousarray((bcs == 1).T).ravel())[0])
        sim.set_electrodes([1e-9, 0.0], [e1.area, e2.area],
                           circuit=dict(R=1.0, L=1e-6, C=1e-6, v0=0.0, v1=1.0, f=5e6, ph=0.0,
                                        pos_ref=ref_node, pos_dof=0, neg_phi=0.0))
        for it in range(24):
            ctx.step(dt, 1)
            sim.advance(dt)
            rho = ctx.field(L.F_RHO)
            assert np.array_equal(rho.view(np.uint64), sim.rho.view(np.uint64)), it
            for k, key in ((0, "driven"), (1, "floating")):
                es = ctx.electrode_state(ids[key])
                assert es["sigma"] == sim.el_sigma[k] and es["dq"] == sim.el_dq[k], (it, key)
            cs = ctx.circuit_state()
            assert (cs["i"], cs["q"], cs["t"], cs["vext"]) == tuple(sim.circ), it
            phi = ctx.field(L.F_PHI)
            A, b, rdof, _ = S.electrode_system(nx, ny, dh, dh, fixed=[(bcs == 3, 0.0)],
                                               floating=[(bcs == 1, sim.el_sigma[0]),
                                                         (bcs == 2, sim.el_sigma[1])])
            ref = S.electrode_potential(A, b, rdof, rho, EPS0)
            assert np.max(np.abs(phi - ref)) <= 1e-8 * np.max(np.abs(ref)), it
            sim.set_phi(phi)
            assert np.array_equal(ctx.field(L.F_EX), sim.E[0::2].astype(np.float64)), it
            sim.end_step(dt)
            if it % 4 == 3:
                stt = ctx.get_state(s)
                assert len(stt["cell"]) == sp.np, it
                for f in ("cell", "ox", "oy", "vx", "vy", "vz", "id"):
                    assert np.array_equal(stt[f].view(np.uint32), getattr(sp, f)[:sp.np].view(
                        np.uint32)), (it, f)
        absorbed = n - sp.np
        # the walls took electrons; the driven electrode's dq is reset every
        # step by the circuit (foo!(::PlasmaDevice)), its sigma moved
        assert absorbed > 0 and sim.el_sigma[0] != 1e-9 and sim.el_dq[0] == 0.0
    finally:
        ctx.close()


def test_host_layer_electrodes_and_circuit_equal_direct_calls(gpu_ready):
    """problem/06_circuit.jl written with the reference-named layer
    (create_electrode, rlc / netlist) against the same run made with direct
    C-ABI calls: electron states and the circuit equal bit for bit."""
    from iskra_amd import _lib as L, circuit as cir
    from iskra_amd import finite_difference_method as fdm, particle_in_cell as pic
    from iskra_amd import regular_grids as rg
    nx = ny = 11
    dh, dt, steps = 0.05, 1e-8, 12
    xs = np.arange(nx) * dh
    grid = rg.create_uniform_grid(xs, xs)
    solver = fdm.create_poisson_solver(grid, EPS0, kind=L.POISSON_PCG, rtol=1e-10)
    e = pic.create_kinetic_species("e-", 2000, -QE, ME, 5e4)
    cfg = pic.Config(grid=grid, solver=solver, pusher=pic.create_boris_pusher(), species=[e])
    bcs = np.zeros((nx, ny), np.int8)
    bcs[0, :] = 1
    bcs[-1, :] = 2
    driven = pic.create_electrode(bcs == 1, cfg, sigma=EPS0)
    grounded = pic.create_electrode(bcs == 2, cfg, fixed=True)
    cfg.circuit = cir.rlc(cir.netlist([
        ("V1", 3, "GND", cir.SineVoltage(1.0, 5e6)),
        ("L1", "NOD", "VCC", 1000e-9),
        ("C1", "NOD", "VCC", 1000e-9),
        ("R1", "GND", "NOD", 1.0),
        ("EXT1", "NOD", "GND", cir.PlasmaDevice(driven, grounded))]))
    src = pic.MaxwellianSource(1e3 / dt, [nx * dh * 0.9, ny * dh * 0.9], [1e5] * 3)
    ctx = pic.bind(cfg)
    try:
        pic.init(src, e, dt)
        ctx.step(dt, steps)
        st_host = ctx.get_state(e._sid)
        c_host = cfg.circuit.state()
        s_host = driven.state()
    finally:
        ctx.close()
    from iskra_amd import Context
    c2 = Context()
    try:
        c2.grid(2, (nx, ny), (dh, dh))
        c2.rng(0)
        c2.poisson(L.POISSON_PCG, EPS0, None, rtol=1e-10)
        flat = lambda m: np.ascontiguousarray(m.T).ravel()  # noqa: E731
        sid = c2.species(-QE, ME, 5e4, 2000, track_ids=True)
        c2.boundary(sid, [L.BND_WRAP, L.BND_WRAP])
        d = c2.electrode(flat(bcs == 1), sigma=EPS0)
        g = c2.electrode(flat(bcs == 2), fixed=True)
        c2.circuit(1.0, 1000e-9, 1000e-9, 0.0, 1.0, 5e6, 0.0, positive=d, negative=g)
        c2.sample_maxwellian(sid, 1000, [nx * dh * 0.9, ny * dh * 0.9], [1e5] * 3,
                             seed=0, stream=(sid << 24) | 0)
        c2.step(dt, steps)
        st2 = c2.get_state(sid)
        for f in ("cell", "ox", "oy", "vx", "vy", "vz"):
            assert np.array_equal(st_host[f], st2[f]), f
        # the reference-named layer numbers particles from 1 (particle_uuids,
        # kinetic.jl:15), the direct calls from 0
        assert np.array_equal(st_host["id"], st2["id"] + 1)
        assert c2.circuit_state() == c_host
        assert c2.electrode_state(d) == s_host
        assert c_host["i"] != 0.0
    finally:
        c2.close()
