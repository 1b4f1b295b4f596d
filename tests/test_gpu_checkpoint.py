"""Checkpoint / resume (iskra_checkpoint_*; no reference counterpart, its
TODO:3): a run interrupted, saved, loaded into a freshly set-up context and
continued ends bit for bit where the uninterrupted run ends (particles
compared by id, since a load re-sorts them by cell; fields and collision
counts directly)."""
import numpy as np
import pytest

from test_gpu_openpmd import _discharge

pytestmark = pytest.mark.gpu


@pytest.fixture
def pic_module():
    import iskra_amd.particle_in_cell as pic
    saved = (pic.enter_loop, pic.after_loop, pic.exit_loop, pic.after_push)
    yield pic
    pic.enter_loop, pic.after_loop, pic.exit_loop, pic.after_push = saved


def _by_id(st):
    keep = st["cell"] >= 0
    o = np.argsort(st["id"][keep], kind="stable")
    return {k: v[keep][o] for k, v in st.items()}


def _assert_same_particles(ca, sa, cb, sb, what):
    a, b = _by_id(ca.get_state(sa)), _by_id(cb.get_state(sb))
    assert len(a["id"]) == len(b["id"]), what
    for k in ("id", "cell", "ox", "oy", "vx", "vy", "vz"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (what, k)


def test_discharge_resumed_from_a_checkpoint_equals_the_whole_run(gpu_ready, pic_module, tmp_path):
    """problem/11_rf_discharge.jl's species and collisions (ionisation makes
    new particles with new ids) through the reference-named layer: 24 steps
    in one go against 10 steps, checkpoint(), and 14 steps of solve(...,
    resume=) on a configuration built anew."""
    pic = pic_module
    from iskra_amd import _lib as L
    dt = None
    a = _discharge(pic)
    dt = a["p"]["dt"]
    print("checkpoint: whole run of 24 steps", flush=True)
    pic.solve(a["config"], dt, 24)
    b = _discharge(pic)
    pic.solve(b["config"], dt, 10)
    path = str(tmp_path / "ck.npz")
    pic.checkpoint(b["config"], path)
    c = _discharge(pic)
    print("checkpoint: resumed run of 14 steps", flush=True)
    pic.solve(c["config"], dt, 14, resume=path)
    ca, cc = a["config"].ctx, c["config"].ctx
    for sa, sc, name in ((a["e"], c["e"], "e-"), (a["ion"], c["ion"], "Ar+")):
        _assert_same_particles(ca, sa._sid, cc, sc._sid, name)
    for f in (L.F_PHI, L.F_RHO, L.F_EX):
        assert np.array_equal(ca.field(f), cc.field(f)), f
    assert np.array_equal(a["electron"].counts(), c["electron"].counts())
    assert np.array_equal(a["ionmcc"].counts(), c["ionmcc"].counts())
    assert a["electron"].counts()[2].sum() > 0  # ionisation happened
    # the species' host mirrors read the resumed device state
    assert c["e"].np == a["e"].np


def _dsmc_ctx(seed=7):
    from iskra_amd import Context, _lib as L
    nodes, dx, n = 33, 1e-3, 8 * 32 * 32
    ctx = Context()
    ctx.grid(2, (nodes, nodes), (dx, dx))
    ctx.tuning(tile=8, chunk=1024, margin=2, sort_every=4)
    ctx.rng(seed)
    mo = 2.6567e-26
    for k, (m, vth) in enumerate(((mo, 3e3), (2 * mo, 1e3))):
        s = ctx.species(0.0, m, 4e17, n, track_ids=True)
        ctx.boundary(s, [L.BND_DISCARD, L.BND_DISCARD])
        ctx.sample_maxwellian(s, n, [32 * dx, 32 * dx], [vth] * 3, seed=seed, stream=k)
    tags = np.zeros((nodes, nodes), np.int8)
    tags[0, :] = tags[-1, :] = tags[:, 0] = tags[:, -1] = 1
    ctx.poisson(L.POISSON_PCG, 8.854e-12, tags.ravel(), rtol=1e-10)
    ctx.tag_value(1, 0.0)
    did = ctx.dsmc(0, 1, [1e2, 1e3, 1e4], [3e-19, 4e-19, 2e-19])
    return ctx, did


def test_dsmc_resumed_through_the_c_abi(gpu_ready):
    """Two gases colliding by DSMC (per-cell remainders carried between
    steps, draws keyed by cell and call): 12 steps against 6 + load + 6."""
    from iskra_amd import _lib as L
    dt = 2e-8
    a, da = _dsmc_ctx()
    a.step(dt, 12)
    b, _ = _dsmc_ctx()
    b.step(dt, 6)
    blob = b.checkpoint()
    b.close()
    c, dc = _dsmc_ctx()
    c.restore(blob)
    c.step(dt, 6)
    for s in (0, 1):
        _assert_same_particles(a, s, c, s, f"species {s}")
    assert np.array_equal(a.field(L.F_PHI), c.field(L.F_PHI))
    assert np.array_equal(a.dsmc_counts(da), c.dsmc_counts(dc))
    assert a.dsmc_counts(da).sum() > 1000
    a.close()
    c.close()


def test_checkpoint_refused_by_another_set_up(gpu_ready):
    from iskra_amd import Context, _lib as L
    a, _ = _dsmc_ctx()
    blob = a.checkpoint()
    other = Context()
    other.grid(2, (17, 17), (1e-3, 1e-3))
    with pytest.raises(L.IskraError) as e:
        other.restore(blob)
    assert e.value.code == L.E_STATE
    with pytest.raises(L.IskraError) as e:
        a.restore(blob[:100])
    assert e.value.code in (L.E_ARG, L.E_STATE)
    a.close()
    other.close()


def _circuit_ctx(orc):
    """problem/07_boundaries.jl's box (driven and floating electrodes, a
    grounded one, reflecting walls) with an RLC circuit and an electron beam,
    as tests/test_gpu_surfaces.py sets it up"""
    from iskra_amd import _lib as L
    from test_gpu_surfaces import ME, QE, _geometry_ctx
    ctx, _, _, _, ids = _geometry_ctx(orc)
    ctx.circuit(R=1.0, L=1e-6, C_=1e-6, v_offset=0.0, v_amplitude=1.0, v_freq=5e6,
                positive=ids["driven"], negative=ids["grounded"])
    s = ctx.species(-QE, ME, 5e4, 4000, track_ids=True)
    ctx.boundary(s, [L.BND_WRAP, L.BND_WRAP])
    ctx.sample_maxwellian(s, 4000, [0.5, 0.5], [3e6] * 3, dx=[0.25, 0.25], seed=3, stream=0)
    return ctx, s, ids


def test_electrodes_and_circuit_resumed(gpu_ready, orc):
    """Surface charges, the circuit's current, charge and time, the driven
    electrode's value and the BiCGStab potential carry over a load: 24 steps
    against 12 + load + 12."""
    from iskra_amd import _lib as L
    dt = 1e-8
    a, sa, ids = _circuit_ctx(orc)
    a.step(dt, 24)
    b, _, _ = _circuit_ctx(orc)
    b.step(dt, 12)
    blob = b.checkpoint()
    b.close()
    c, sc, _ = _circuit_ctx(orc)
    c.restore(blob)
    c.step(dt, 12)
    _assert_same_particles(a, sa, c, sc, "electrons")
    assert np.array_equal(a.field(L.F_PHI), c.field(L.F_PHI))
    assert a.circuit_state() == c.circuit_state()
    for k in ("driven", "floating"):
        assert a.electrode_state(ids[k]) == c.electrode_state(ids[k]), k
    assert a.circuit_state()["i"] != 0.0
    a.close()
    c.close()
