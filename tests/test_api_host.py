"""Host-side behaviour of the reference-named API (no GPU needed): grids,
cell volumes, solver bookkeeping, reaction parsing and the boundary hook
rules.  Every particle or field operation itself runs on the GPU and is
tested in test_gpu_api.py."""
import numpy as np
import pytest

from iskra_amd import chemistry as ch
from iskra_amd import finite_difference_method as fdm
from iskra_amd import particle_in_cell as pic
from iskra_amd import regular_grids as rg
from iskra_amd import _lib as L


def test_create_uniform_grid_matches_reference_convention():
    # RegularGrids.jl:55-69: n cells -> n+1 nodes from a range, dx from its step
    g = rg.create_uniform_grid(np.arange(0, 1.01, 0.5), np.arange(0, 2.01, 0.5))
    assert g.size() == (3, 5) and len(g) == 15
    assert g.dh == (0.5, 0.5) and g.origin == (0.0, 0.0)
    assert np.array_equal(g.coords[0][:, 0], [0.0, 0.5, 1.0])
    assert np.array_equal(g.coords[1][0, :], [0.0, 0.5, 1.0, 1.5, 2.0])
    g1 = rg.create_uniform_grid(np.linspace(0, 1, 9))
    assert g1.size() == (9,) and g1.dh == (0.125,)


def test_cell_volume_edges_and_corners():
    # RegularGrids.jl:26-38: dx*dy, halved on non-periodic edges, so corners 1/4
    g = rg.create_uniform_grid(np.arange(4) * 0.5, np.arange(3) * 2.0)
    V = rg.cell_volume(g)
    assert V[1, 1] == 1.0 and V[0, 1] == 0.5 and V[1, 0] == 0.5 and V[0, 0] == 0.25
    assert V[-1, -1] == 0.25
    # total volume of the domain
    assert V.sum() == pytest.approx(1.5 * 4.0)
    gp = rg.create_uniform_grid(np.arange(4) * 0.5, np.arange(3) * 2.0, left="periodic",
                                right="periodic")
    Vp = rg.cell_volume(gp)
    assert Vp[0, 1] == 1.0 and Vp[0, 0] == 0.5
    assert gp.vol_periodic() == (1, 1, 0, 0)


def test_apply_periodic_axis_labels_follow_reference():
    # generalized_poisson.jl:291-323: axis 1 links the second index, axis 2 the first
    g = rg.create_uniform_grid(np.arange(5.0), np.arange(5.0))
    ps = fdm.create_poisson_solver(g, 1.0)
    fdm.apply_periodic(ps, 1)
    assert ps.periodic == [False, True]
    fdm.apply_periodic(ps, 2)
    assert ps.periodic == [True, True]
    with pytest.raises(ValueError):
        fdm.apply_periodic(ps, 3)
    g1 = rg.create_uniform_grid(np.arange(9.0))
    ps1 = fdm.create_poisson_solver(g1, 1.0)
    fdm.apply_periodic(ps1)
    assert ps1.periodic == [True, False]


def test_apply_dirichlet_tags_and_value_updates():
    g = rg.create_uniform_grid(np.arange(0, 1.01, 0.5), np.arange(0, 2.01, 0.5))
    ps = fdm.create_poisson_solver(g, 1.0)
    bcs = np.zeros(g.size(), np.int8)
    bcs[0, :] = 1
    bcs[2, :] = 2
    fdm.apply_dirichlet(ps, bcs == 1, 0.0)
    fdm.apply_dirichlet(ps, bcs == 2, 1.0)
    assert np.array_equal(ps.tags, bcs)
    # the same mask again only changes its value (RF electrode pattern)
    fdm.apply_dirichlet(ps, bcs == 1, 450.0)
    assert len(ps.masks) == 2 and ps.values == [450.0, 1.0]
    # solver choice: two Dirichlet columns only -> PCG; all four walls of a
    # 2^k+1 grid -> multigrid; 1-D -> direct
    assert ps._solver_kind() == L.POISSON_PCG
    g2 = rg.create_uniform_grid(np.arange(17.0), np.arange(17.0))
    ps2 = fdm.create_poisson_solver(g2, 1.0)
    walls = np.zeros(g2.size(), bool)
    walls[0, :] = walls[-1, :] = walls[:, 0] = walls[:, -1] = True
    fdm.apply_dirichlet(ps2, walls, 0.0)
    assert ps2._solver_kind() == L.POISSON_MG
    assert fdm.create_poisson_solver(rg.create_uniform_grid(np.arange(9.0)),
                                     1.0)._solver_kind() == L.POISSON_DIRECT
    B = fdm.calculate_magnetic_field(ps2)
    assert B.shape == (17, 17, 3) and not B.any()


def test_reactions_and_mcc_accept():
    # reactions.jl:14-55 (net stoichiometry) and mcc.jl:291-320 (accept)
    e = pic.create_kinetic_species("e-", 10, -1.0, 1.0, 1.0, D=1)
    iHe = pic.create_kinetic_species("He+", 10, 1.0, 4.0, 1.0, D=1)
    He = pic.FluidSpecies("He", 1.0, 0.0, 4.0, np.ones(5), 300.0)
    s1 = ch.CrossSection([0.0, 10.0], [1e-20, 1e-20])
    s4 = ch.CrossSection(np.array([[24.0, 0.0], [100.0, 1e-20]]))
    rs = ch.reactions([
        (s1, [e, He], [e, He]),
        (s4, [e, He], [e, e, iHe], ch.Ionization(24.587)),
    ])
    assert rs[0].stoichiometry == [] and rs[0].type is None
    assert [(s.name, c) for s, c in rs[1].stoichiometry] == [("e-", 1), ("He", -1), ("He+", 1)]
    m = ch.mcc(rs)
    assert isinstance(m.collisions[0].type, ch.ElasticIsotropic)   # default type
    assert m.collisions[1].products == [e, iHe]
    assert m.source is e and m.target is He
    assert s4.maximum() == 1e-20 and s4.argmax() == 100.0
    with pytest.raises(ValueError, match="fluid"):
        ch.mcc(ch.reactions([(s1, [e, iHe], [e, iHe])]))


def test_boundary_calls_outside_after_push_are_rejected():
    g = rg.create_uniform_grid(np.arange(9.0))
    e = pic.create_kinetic_species("e-", 10, -1.0, 1.0, 1.0, D=1)
    with pytest.raises(RuntimeError, match="after_push"):
        pic.wrap(e, g)
    # inside the hook the calls record the per-dimension boundary
    modes = [L.BND_NONE, L.BND_NONE]
    pic._recording[id(e)] = modes
    try:
        pic.discard(e, g, dims=1)
    finally:
        del pic._recording[id(e)]
    assert modes == [L.BND_DISCARD, L.BND_NONE]


def test_species_fields_and_helpers():
    # create_kinetic_species / thermal_speed / create_thermalized_beam
    # (configuration.jl:79-102; kinetic.jl:15-18)
    e = pic.create_kinetic_species("e-", 4, -2.0, 3.0, 5.0, D=2)
    assert e.x.shape == (4, 2) and e.v.shape == (4, 3)
    assert np.array_equal(e.id, [1, 2, 3, 4]) and np.all(e.wg == 5.0) and e.w0 == 5.0
    assert e.np == 0
    assert pic.thermal_speed(300.0, 1.0) == pytest.approx(np.sqrt(2 * pic.KB * 300.0))
    src = pic.create_thermalized_beam(e, [1.0, 2.0], [0.5, 0.0, 0.0], T=0.0, rate=8.0)
    assert np.array_equal(src.wx, [1.0, 2.0]) and np.array_equal(src.dv, [0.5, 0, 0])
    assert np.all(src.wv == 0.0)
    with pytest.raises(ValueError):
        e.np = 5


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus above 1 exits with status 2 and a
    message, before any device work."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_fluid_classification_and_moving_target_gas_refused():
    """advection_diffusion.jl:4-38 (called by advance!(::FluidSpecies),
    ParticleInCell.jl:74-82) leaves n unchanged for a uniform 1-D fluid, any
    fluid with mu = 0 in 1-D, and a neutral uniform 2-D fluid.  Fluids whose
    parameters let the update change them (mu != 0, or q != 0 in 2-D) are
    advanced on the device, except an MCC target gas, which must stay
    constant and is refused before any device work when it would not."""
    g1 = rg.create_uniform_grid(np.linspace(0, 1, 17))
    g2 = rg.create_uniform_grid(np.linspace(0, 1, 9), np.linspace(0, 1, 9))
    g2q = rg.create_uniform_grid(np.linspace(0, 1, 9), np.linspace(0, 1, 2))
    ramp = np.linspace(1.0, 2.0, 17) ** 2
    static = [(g1, pic.FluidSpecies("bg", 0.0, 1.0, 1.0, ramp, 300.0)),
              (g1, pic.FluidSpecies("bg", 1e-3, 1.0, 1.0, np.full(17, 3.0), 300.0)),
              (g2, pic.FluidSpecies("gas", 1e-3, 0.0, 1.0, np.full((9, 9), 3.0), 300.0)),
              (g2q, pic.FluidSpecies("ions", 1e-3, 1.0, 1.0, np.full((9, 2), 3.0), 300.0))]
    moving = [(g1, pic.FluidSpecies("bg", 1e-3, 1.0, 1.0, ramp, 300.0)),
              (g2, pic.FluidSpecies("ions", 0.0, 1.0, 1.0, np.full((9, 9), 3.0), 300.0)),
              (g2, pic.FluidSpecies("gas", 1e-3, 0.0, 1.0,
                                    np.outer(np.arange(9.0), np.arange(9.0)) ** 2, 300.0))]
    for g, f in static:
        assert pic._fluid_is_static(f, g), f.name
    for g, f in moving:
        assert not pic._fluid_is_static(f, g), f.name
        assert pic._fluid_can_change(f, g), f.name
        e = pic.create_kinetic_species("e-", 10, -1.0, 1.0, 1.0, D=g.dim)
        sig = ch.CrossSection(np.array([0.0, 10.0]), np.array([1e-20, 1e-20]))
        coll = ch.mcc(ch.reactions([(sig, [e, f], [e, f])]))
        cfg = pic.Config(grid=g, species=[e, f], interactions=[coll])
        with pytest.raises(L.IskraError) as ei:
            pic.bind(cfg)
        assert ei.value.code == L.E_UNSUPPORTED and cfg.ctx is None
    assert not pic._fluid_can_change(pic.FluidSpecies("bg", 0.0, 1.0, 1.0, ramp, 300.0), g1)
    assert not pic._fluid_can_change(pic.FluidSpecies("gas", 0.0, 0.0, 1.0, 0.0, 300.0), g2)
    assert pic._fluid_can_change(pic.FluidSpecies("ions", 0.0, 1.0, 1.0, 0.0, 300.0), g2)


def test_axial_grid_pusher_and_solver_host_side(orc):
    """create_axial_grid (RegularGrids.jl:84-97): r fastest node layout,
    radial bcs (:other, :other), node volumes equal to the oracle's
    restatement of cell_volume(::AxialGrid{2}); the axial pusher and the
    PCG default; apply_periodic refused (the reference's RZ version cannot
    run)."""
    g = rg.create_axial_grid(np.arange(0, 9) * 0.5e-3, np.arange(0, 6) * 1e-3, top=rg.PERIODIC)
    assert g.kind == "rz" and g.n == (9, 6) and g.dh == (0.5e-3, 1e-3)
    assert g.bcs == (("other", "other"), ("open", "periodic"))
    V = rg.cell_volume(g)
    want = orc.cell_volume_axial(9, 6, 0.5e-3, 1e-3, (0, 1)).reshape(6, 9).T
    assert np.array_equal(V, want)
    assert pic.create_axial_boris_pusher().kind == "rz"
    assert pic.create_boris_pusher().kind == "xy"
    ps = fdm.create_poisson_solver(g, 8.854e-12)
    assert ps._solver_kind() == L.POISSON_PCG
    with pytest.raises(L.IskraError) as ei:
        fdm.apply_periodic(ps, 1)
    assert ei.value.code == L.E_UNSUPPORTED
