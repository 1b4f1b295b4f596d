"""The HIP path's post-collision states against the reference's collision
kernels (Chemistry/src/mcc.jl:129-229), one collision type at a time, at the
shape of the 1-D RF discharge (513 nodes, 2e5 electrons).

The draws of every particle come from the device through the C ABI
(iskra_rng_export); which particles collided is read from the states
(iskra_species_get_state) before and after one iskra_mcc_run.  For each of
them the test recomputes the collision in numpy from the reference's
formulas -- `unrotated` applied as its matrix literal (a row vector times the
matrix), Box-Muller normals, the reference's angle distributions -- and checks
the FP32 result, plus the closed-form invariants: the scattered relative speed
|v' - w| = mr2 |v_r| |r| with |r| from the closed form of that (non-orthogonal)
matrix, the excitation energy eps' = (eps - E_th) |r|^2, the ionisation energy
split, the secondary's direction relative to the updated primary and the ions'
Maxwellian velocity.  Every collided particle must have been a null-collision
candidate (Uc < N max_P) and their number must equal the events counted."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

QE = 1.602176634e-19
ME = 9.1093837015e-31
MAR = 39.948 * 1.66053906660e-27
QE_MCC = 1.60217646e-19
KB_MCC = 1.3806503e-23
M_EV = ME / QE_MCC
VTH = np.sqrt(2.0 * KB_MCC * 300.0 / MAR)
MR1, MR2 = ME / (ME + MAR), MAR / (ME + MAR)


def u01(w):
    """52-bit uniform in (0,1) from two Philox words (the library's mapping)."""
    a = w[..., 0].astype(np.uint64) >> np.uint64(6)
    b = w[..., 1].astype(np.uint64) >> np.uint64(6)
    return (((a << np.uint64(26)) | b).astype(np.float64) + 0.5) / 4503599627370496.0


def uniforms(words):
    """(n, blocks, 4) Philox words -> (n, 2 blocks) uniforms in draw order."""
    n, nb, _ = words.shape
    u = np.empty((n, 2 * nb))
    u[:, 0::2] = u01(words[:, :, 0:2])
    u[:, 1::2] = u01(words[:, :, 2:4])
    return u


def euler(v):
    nv = np.sqrt((v * v).sum(1))
    ct = v[:, 2] / nv
    st = np.sqrt(1.0 - ct * ct)
    pole = st == 0.0
    sst = np.where(pole, 1.0, st)
    cp = np.where(pole, 1.0, v[:, 0] / nv / sst)
    sp = np.where(pole, 0.0, v[:, 1] / nv / sst)
    return st, ct, sp, cp


def scatter(v, kind, u1, u2):
    """vec([sinX cosE sinX sinE cosX] * unrotated(euler_angles(v)))
    (mcc.jl:53-127); kind 0 "isotropic" (chi = 2 pi U), 1 diffuse."""
    st, ct, sp, cp = euler(v)
    if kind == 0:
        sx, cx = np.sin(2 * np.pi * u1), np.cos(2 * np.pi * u1)
    else:
        sx = np.sqrt(u1)
        cx = -np.sqrt(1.0 - sx * sx)
    se, ce = np.sin(2 * np.pi * u2), np.cos(2 * np.pi * u2)
    T = np.array([[cp * ct, -sp * ct, -st],
                  [-sp, cp, np.zeros_like(st)],
                  [cp * st, sp * st, ct]])  # (3, 3, n): the literal of mcc.jl:83-87
    a = np.array([sx * ce, sx * se, cx])
    r = np.einsum("in,ijn->nj", a, T)
    a0, a1, a2 = a
    r2 = 1.0 - 4.0 * a0 * sp * ct * (a1 * cp + a2 * sp * st)
    return r, r2


def randn3(u):
    r1, r2 = np.sqrt(-2.0 * np.log(u[:, 0])), np.sqrt(-2.0 * np.log(u[:, 2]))
    return np.stack([r1 * np.cos(2 * np.pi * u[:, 1]), r1 * np.sin(2 * np.pi * u[:, 1]),
                     r2 * np.cos(2 * np.pi * u[:, 3])], axis=1)


def _setup(type_, ne=200_000, seed=2024):
    from iskra_amd import Context, _lib as L
    nodes, gap = 513, 0.025
    dx = gap / (nodes - 1)
    rng = np.random.default_rng(seed)
    ctx = Context()
    ctx.grid(1, (nodes,), (dx,))
    ctx.rng(seed)
    e = ctx.species(-QE, ME, 1e8, 3 * ne, track_ids=True)
    i = ctx.species(QE, MAR, 1e8, 3 * ne, track_ids=True)
    ctx.boundary(e, [L.BND_DISCARD])
    ctx.boundary(i, [L.BND_DISCARD])
    v = rng.standard_normal((ne, 3)) * 2.2e6  # ~14 eV per component
    ctx.upload(e, rng.random(ne)[:, None] * gap, v)
    energy = {2: 11.5, 3: 15.8}.get(type_, 0.0)
    if energy:
        # zero at threshold (A18), constant above
        eps, sig = [energy, energy + 1.0, 1000.0], [0.0, 1e-19, 1e-19]
    else:
        eps, sig = [0.0, 1000.0], [2e-19, 2e-19]
    mid = ctx.mcc(e, [type_], [energy], [(eps, sig)], ion=i if type_ == 3 else -1,
                  target_m=MAR, target_q=0.0, target_T=300.0)
    ng = 3.219e22
    ctx.set_field(L.F_TARGET, np.full(nodes, ng), index=mid)
    # candidate probability as the library forms it (mcc.jl:27-51,242-248)
    e_nodes = np.unique(np.concatenate([[0.0], eps]))
    sg = np.interp(e_nodes, eps, sig) * np.sqrt(2.0 * e_nodes / M_EV)
    dt = 2e-12
    candP = 1.0 - np.exp(-ng * sg.max() * dt)
    return ctx, e, i, mid, energy, dt, candP


@pytest.mark.parametrize("type_", [0, 1, 2, 3, 4])
def test_collision_kinematics_on_device(gpu_ready, type_):
    ctx, e, i, mid, energy, dt, candP = _setup(type_)
    try:
        print(f"collision type {type_}", flush=True)
        before = ctx.get_state(e)
        n = len(before["cell"])
        ev = ctx.mcc_run(mid, dt)
        after = ctx.get_state(e)
        ions = ctx.get_state(i)
        u = uniforms(ctx.rng_export(0x4D430000 + mid, 0, before["id"], 6))
        v0 = np.stack([before["vx"], before["vy"], before["vz"]], 1).astype(np.float64)
        v1 = np.stack([after["vx"][:n], after["vy"][:n], after["vz"][:n]], 1)
        hit = np.nonzero(np.any(v1 != v0.astype(np.float32), axis=1))[0]
        assert 1000 < len(hit) < n // 2, len(hit)
        assert len(hit) == ev[0]
        assert np.all(u[hit, 0] < candP * (1 + 1e-12))
        vin, uu = v0[hit], u[hit]
        got = v1[hit].astype(np.float64)
        if type_ in (0, 1, 4):
            tv = VTH * randn3(uu[:, 2:6])
            vr = vin - tv
            w = MR1 * vin + MR2 * tv
            nvr = np.sqrt((vr * vr).sum(1))
            if type_ == 4:
                s, ua, ub = uu[:, 6], uu[:, 7], uu[:, 8]
            else:
                s, ua, ub = 1.0, uu[:, 6], uu[:, 7]
            r, r2 = scatter(vr, 0 if type_ == 0 else 1, ua, ub)
            want = w + (MR2 * (s * nvr))[:, None] * r
            # invariant: scattered relative speed (|w| ~ |v|: FP32 rounding of v')
            d = np.sqrt(((got - w) ** 2).sum(1))
            ref = MR2 * s * nvr * np.sqrt(r2)
            assert np.all(np.abs(d - ref) <= 2e-7 * np.sqrt((got * got).sum(1)) + 1e-6 * ref)
        elif type_ == 2:
            eps = 0.5 * M_EV * (vin * vin).sum(1)
            assert np.all(eps >= energy)
            r, r2 = scatter(vin, 0, uu[:, 2], uu[:, 3])
            want = (np.sqrt(2.0 / M_EV) * np.sqrt(eps - energy))[:, None] * r
            eps1 = 0.5 * M_EV * (got * got).sum(1)
            assert np.allclose(eps1, (eps - energy) * r2, rtol=1e-6, atol=1e-9 * energy)
        else:
            eps = 0.5 * M_EV * (vin * vin).sum(1)
            sE = eps - energy
            assert np.all(sE >= 0)
            al = np.sqrt(2.0 / M_EV)
            e1 = sE * uu[:, 2]
            e2 = sE - e1
            r, r2 = scatter(vin, 1, uu[:, 3], uu[:, 4])
            want = (al * np.sqrt(e1))[:, None] * r
            r_2, r2_2 = scatter(want, 1, uu[:, 5], uu[:, 6])
            sec = (al * np.sqrt(e2))[:, None] * r_2
            vion = VTH * randn3(uu[:, 7:11])
            eps1 = 0.5 * M_EV * (got * got).sum(1)
            assert np.allclose(eps1, e1 * r2, rtol=1e-6, atol=1e-9 * energy)
            # products: one electron on the source species and round(w0/w0) = 1
            # ion per event, both at the parent's position
            npro = len(after["cell"]) - n
            assert npro == len(hit) and len(ions["cell"]) == len(hit)
            key = lambda c, o: c.astype(np.int64) * 2**32 + o.view(np.uint32)
            kp = key(before["cell"][hit], before["ox"][hit])
            assert len(np.unique(kp)) == len(kp)
            order = np.argsort(kp)
            ke = key(after["cell"][n:], after["ox"][n:])
            oe = np.argsort(ke)
            assert np.array_equal(ke[oe], kp[order])
            gs = np.stack([after["vx"][n:], after["vy"][n:], after["vz"][n:]], 1)[oe]
            ws = sec[order]
            assert np.allclose(gs, ws, rtol=0, atol=4e-7 * np.abs(ws).max(axis=1,
                                                                       keepdims=True) + 1e-3)
            eps2 = 0.5 * M_EV * (gs.astype(np.float64) ** 2).sum(1)
            assert np.allclose(eps2, e2[order] * r2_2[order], rtol=2e-6, atol=1e-6 * energy)
            ki = key(ions["cell"], ions["ox"])
            oi = np.argsort(ki)
            assert np.array_equal(ki[oi], kp[order])
            gi = np.stack([ions["vx"], ions["vy"], ions["vz"]], 1)[oi]
            assert np.allclose(gi, vion[order], rtol=1e-6, atol=1e-4)
        # the post-collision velocity itself (FP32 rounding of the FP64 result)
        tol = 3e-7 * np.abs(want).max(axis=1, keepdims=True) + 1e-3
        assert np.all(np.abs(got - want) <= tol), np.abs(got - want).max()
    finally:
        ctx.close()
