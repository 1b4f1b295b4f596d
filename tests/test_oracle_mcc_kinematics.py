"""Pins the oracle's MCC collision kinematics (mir_scatter, mir_coll_kin) to
the reference's collision kernels, Chemistry/src/mcc.jl:53-229, without
retyping the oracle's arithmetic:

* the scattering direction is checked against the reference's `unrotated`
  matrix literal (mcc.jl:83-87) applied as the reference applies it
  (`vec([sinX cosE sinX sinE cosX] * T)`, a row vector times the matrix),
  against the incident direction at chi = 0, and against closed forms of its
  length and of its angle to the incident velocity;
* the angles come from the reference's distributions (mcc.jl:53-72): chi
  uniform on [0, 2 pi) for "isotropic", sin chi = sqrt(U) with cos chi < 0 for
  diffuse reflection (A12), checked pointwise and by KS tests;
* each collision type is checked by its physics with the draws injected:
  Box-Muller target velocity, fixed centre-of-mass velocity, the scattered
  relative speed, energy conservation, excitation energy loss, the
  ionisation energy split and the secondary's direction relative to the
  updated primary (A17).

The reference's `unrotated` matrix is not orthogonal: Julia reads the first
row `[ cosp*cost -sinp*cost -sint ...` as three elements, so its (1,2) entry
is -sin(phi) cos(theta) where a rotation has +sin(phi) cos(theta).  Scattered
directions are therefore unit vectors only when cos(eta) = 0 or the velocity
lies in the x-z plane; elsewhere the reference's collisions scale |v_r| (and
the kinetic energy after excitation/ionisation) by |r|^2 =
1 - 4 a0 sin(phi) cos(theta) (a1 cos(phi) + a2 sin(phi) sin(theta)) with
a = (sin chi cos eta, sin chi sin eta, cos chi).  The build reproduces this,
and the tests below pin it by that closed form, and pin the textbook
invariants exactly where |r| = 1.
"""
import numpy as np
import pytest
from scipy import stats

import oracle as O

ME = 9.1093837015e-31
MAR = 39.948 * 1.66053906660e-27
M_EV = ME / O.QE_MCC
VTH_AR = np.sqrt(2.0 * O.KB_MCC * 300.0 / MAR)


def euler_angles(v):
    """euler_angles (mcc.jl:89-106); `sin(theta) ~ 0.0` is Julia's isapprox
    against zero with no absolute tolerance, i.e. exact zero."""
    nv = np.sqrt(v @ v)
    ct = v[2] / nv
    st = np.sqrt(1.0 - ct * ct)
    if st == 0.0:
        return st, ct, 0.0, 1.0
    return st, ct, v[1] / nv / st, v[0] / nv / st


def unrotated_literal(st, ct, sp, cp):
    """The matrix literal of mcc.jl:83-87, element by element as Julia parses
    it (whitespace separates elements, so `-sinp*cost` is its own entry)."""
    return np.array([[cp * ct, -sp * ct, -st],
                     [-sp, cp, 0.0],
                     [cp * st, sp * st, ct]])


def angles(kind, u1, u2):
    """(sin chi, cos chi, sin eta, cos eta) from the reference's
    isotropic_distribution (kind 0) or cosine_distribution (kind 1)."""
    if kind == 0:
        chi = 2.0 * np.pi * u1
        sx, cx = np.sin(chi), np.cos(chi)
    else:
        sx = np.sqrt(u1)
        cx = -np.sqrt(1.0 - sx * sx)
    eta = 2.0 * np.pi * u2
    return sx, cx, np.sin(eta), np.cos(eta)


def closed_forms(v, kind, u1, u2):
    """|r|^2 and r . v_hat of the reference's scattered direction, from the
    Gram matrix of its rows: G01 = -2 sp cp ct, G02 = -2 sp^2 ct st, G12 = 0."""
    st, ct, sp, cp = euler_angles(v)
    sx, cx, se, ce = angles(kind, u1, u2)
    a0, a1, a2 = sx * ce, sx * se, cx
    r2 = 1.0 - 4.0 * a0 * sp * ct * (a1 * cp + a2 * sp * st)
    cosang = cx - 2.0 * a0 * sp * sp * ct * st
    return r2, cosang


def cond(v):
    """Round-off amplification of the angles of v: sin(theta) = sqrt(1-cos^2)
    loses eps/sin^2(theta), and cos(phi), sin(phi) divide by sin(theta)."""
    st = max(euler_angles(v)[0], 1e-3)
    return 1.0 + 1.0 / (st * st)


def box_muller(u):
    r1, r2 = np.sqrt(-2.0 * np.log(u[0])), np.sqrt(-2.0 * np.log(u[2]))
    return np.array([r1 * np.cos(2 * np.pi * u[1]), r1 * np.sin(2 * np.pi * u[1]),
                     r2 * np.cos(2 * np.pi * u[3])])


def _rand_v(rng, scale=3e6):
    return rng.standard_normal(3) * scale


# --------------------------------------------------------------------------
# scattering direction
# --------------------------------------------------------------------------
def test_scatter_is_the_reference_matrix_product(orc):
    rng = np.random.default_rng(1)
    vs = [_rand_v(rng) for _ in range(300)]
    vs += [np.array([0.0, 0.0, 2.5e6]), np.array([0.0, 0.0, -1e5]),
           np.array([3e5, 0.0, 0.0]), np.array([0.0, -4e5, 1e5])]
    for v in vs:
        for kind in (0, 1):
            u1, u2 = rng.random(), rng.random()
            sx, cx, se, ce = angles(kind, u1, u2)
            want = np.array([sx * ce, sx * se, cx]) @ unrotated_literal(*euler_angles(v))
            got = O.scatter(kind, v, u1, u2)
            # |v| is summed in another order by numpy's dot
            tol = 2e-15 * cond(v)
            assert np.allclose(got, want, rtol=0, atol=tol), (v, kind, got, want)


def test_scatter_chi_zero_and_pi_follow_the_incident_direction(orc):
    rng = np.random.default_rng(2)
    for _ in range(100):
        v = _rand_v(rng)
        vh = v / np.linalg.norm(v)
        u2 = rng.random()
        assert np.allclose(O.scatter(0, v, 0.0, u2), vh, rtol=0, atol=1e-15)
        assert np.allclose(O.scatter(0, v, 0.5, u2), -vh, rtol=0, atol=1e-15)
        # diffuse with sin chi = 0 is straight back
        assert np.allclose(O.scatter(1, v, 0.0, u2), -vh, rtol=0, atol=1e-15)


def test_reference_matrix_is_not_orthogonal_and_closed_forms_hold(orc):
    rng = np.random.default_rng(3)
    dev = []
    for _ in range(2000):
        v = _rand_v(rng)
        vh = v / np.linalg.norm(v)
        for kind in (0, 1):
            u1, u2 = rng.random(), rng.random()
            r = O.scatter(kind, v, u1, u2)
            r2, cosang = closed_forms(v, kind, u1, u2)
            tol = 1e-14 * cond(v)
            assert r @ r == pytest.approx(r2, rel=0, abs=tol)
            assert r @ vh == pytest.approx(cosang, rel=0, abs=tol)
            dev.append(abs(r @ r - 1.0))
            # where cos(eta) = 0 the matrix acts as a rotation: unit length and
            # the angle to the incident direction is chi
            for u2q in (0.25, 0.75):
                rq = O.scatter(kind, v, u1, u2q)
                sx, cx, _, _ = angles(kind, u1, u2q)
                assert rq @ rq == pytest.approx(1.0, abs=tol)
                assert rq @ vh == pytest.approx(cx, abs=tol)
    # the reference's quirk is real, not round-off: lengths off by tens of %
    assert max(dev) > 0.3


def test_scatter_angle_distributions(orc):
    """chi uniform on [0, 2 pi) (the reference's "isotropic", A12): cos chi has
    the arcsine law P(c <= x) = 1 - acos(x)/pi; diffuse: cos chi = -sqrt(1-U),
    P(c <= x) = 1 - x^2 on [-1, 0].  Measured as the angle to the incident
    direction where cos(eta) = 0."""
    rng = np.random.default_rng(4)
    n = 20000
    u1 = rng.random(n)
    iso, dif = np.empty(n), np.empty(n)
    for q in range(n):
        v = _rand_v(rng)
        vh = v / np.linalg.norm(v)
        iso[q] = O.scatter(0, v, u1[q], 0.25) @ vh
        dif[q] = O.scatter(1, v, u1[q], 0.25) @ vh
    arcsine = lambda x: 1.0 - np.arccos(np.clip(x, -1, 1)) / np.pi
    lambert = lambda x: np.where(x < -1, 0.0, np.where(x > 0, 1.0, 1.0 - x * x))
    assert stats.kstest(iso, arcsine).pvalue > 1e-3
    assert stats.kstest(dif, lambert).pvalue > 1e-3
    assert dif.max() <= 0.0
    # the tests have power: cos chi uniform on [-1, 1] (isotropic on the
    # sphere) and a forward hemisphere are rejected
    assert stats.kstest(2 * rng.random(n) - 1, arcsine).pvalue < 1e-6
    assert stats.kstest(-dif, lambert).pvalue < 1e-6


# --------------------------------------------------------------------------
# collision kernels with injected draws
# --------------------------------------------------------------------------
@pytest.mark.parametrize("type_", [0, 1, 4])
def test_elastic_kinematics(orc, type_):
    """ElasticIsotropic / ElasticBackward / InelasticBackward (mcc.jl:129-172):
    tv = vth randn(3); w = mr1 v + mr2 tv is kept; v' - w = mr2 s |v_r| r with
    s = 1, or the extra uniform of InelasticBackward."""
    rng = np.random.default_rng(10 + type_)
    mr1, mr2 = ME / (ME + MAR), MAR / (ME + MAR)
    kind = 0 if type_ == 0 else 1
    for _ in range(500):
        v = _rand_v(rng)
        u = rng.random(10)
        used, v1, tv, _ = O.coll_kin(type_, ME, MAR, M_EV, 0.0, VTH_AR, v, u)
        assert used == (7 if type_ == 4 else 6)
        assert np.allclose(tv, VTH_AR * box_muller(u[:4]), rtol=1e-13, atol=1e-14 * VTH_AR)
        vr = v - tv
        w = mr1 * v + mr2 * tv
        s, ua, ub = (u[4], u[5], u[6]) if type_ == 4 else (1.0, u[4], u[5])
        r2, cosang = closed_forms(vr, kind, ua, ub)
        d = v1 - w
        nd = np.linalg.norm(d)
        assert nd == pytest.approx(mr2 * s * np.linalg.norm(vr) * np.sqrt(r2), rel=1e-11)
        assert d @ vr / (nd * np.linalg.norm(vr)) == pytest.approx(cosang / np.sqrt(r2),
                                                                    abs=1e-11)
    # where |r| = 1 an elastic collision conserves the pair's momentum and
    # energy (the target's velocity after it is w - mr1 |v_r| r)
    if type_ != 4:
        for _ in range(200):
            v = _rand_v(rng)
            u = rng.random(10)
            u[5] = 0.25
            _, v1, tv, _ = O.coll_kin(type_, ME, MAR, M_EV, 0.0, VTH_AR, v, u)
            w = mr1 * v + mr2 * tv
            t1 = w - (mr1 / mr2) * (v1 - w)
            p0, p1 = ME * v + MAR * tv, ME * v1 + MAR * t1
            assert np.allclose(p1, p0, rtol=1e-12, atol=1e-12 * np.abs(p0).max())
            k0 = ME * (v @ v) + MAR * (tv @ tv)
            k1 = ME * (v1 @ v1) + MAR * (t1 @ t1)
            assert k1 == pytest.approx(k0, rel=1e-11)
            # and the relative speed is unchanged
            assert np.linalg.norm(v1 - t1) == pytest.approx(np.linalg.norm(v - tv), rel=1e-11)


def test_excitation_energy_loss(orc):
    """Excitation (mcc.jl:215-229): eps' = (eps - E_th) |r|^2 along the
    reference's "isotropic" direction; eps' = eps - E_th exactly where
    |r| = 1; below threshold nothing happens."""
    rng = np.random.default_rng(20)
    Eth = 11.5
    for _ in range(500):
        v = _rand_v(rng, 4e6)
        eps = 0.5 * M_EV * (v @ v)
        u = rng.random(10)
        used, v1, _, _ = O.coll_kin(2, ME, MAR, M_EV, Eth, VTH_AR, v, u)
        if eps < Eth:
            assert used == 0
            continue
        assert used == 2
        r2, cosang = closed_forms(v, 0, u[0], u[1])
        eps1 = 0.5 * M_EV * (v1 @ v1)
        assert eps1 == pytest.approx((eps - Eth) * r2, rel=1e-12)
        assert v1 @ v / (np.linalg.norm(v1) * np.linalg.norm(v)) == pytest.approx(
            cosang / np.sqrt(r2), abs=1e-12)
        u[1] = 0.75
        _, v1, _, _ = O.coll_kin(2, ME, MAR, M_EV, Eth, VTH_AR, v, u)
        assert 0.5 * M_EV * (v1 @ v1) == pytest.approx(eps - Eth, rel=1e-12)


def test_ionisation_energy_split_and_secondary(orc):
    """Ionization (mcc.jl:174-213): e1 = (eps - E_iz) U, e2 = rest; the primary
    leaves along a diffuse direction from its old velocity, the secondary along
    a diffuse direction from the primary's NEW velocity (A17); the ions get
    vth randn(3)."""
    rng = np.random.default_rng(30)
    Eiz = 15.8
    seen = 0
    for _ in range(800):
        v = _rand_v(rng, 5e6)
        eps = 0.5 * M_EV * (v @ v)
        u = rng.random(10)
        used, v1, vion, v2 = O.coll_kin(3, ME, MAR, M_EV, Eiz, VTH_AR, v, u)
        if eps < Eiz:
            assert used == 0
            continue
        seen += 1
        assert used == 9
        sE = eps - Eiz
        r2a, ca = closed_forms(v, 1, u[1], u[2])
        r2b, cb = closed_forms(v1, 1, u[3], u[4])
        e1 = 0.5 * M_EV * (v1 @ v1)
        e2 = 0.5 * M_EV * (v2 @ v2)
        assert e1 == pytest.approx(sE * u[0] * r2a, rel=1e-11)
        assert e2 == pytest.approx(sE * (1.0 - u[0]) * r2b, rel=1e-9, abs=1e-12 * sE)
        assert v1 @ v / (np.linalg.norm(v1) * np.linalg.norm(v)) == pytest.approx(
            ca / np.sqrt(r2a), abs=1e-11)
        if e2 > 1e-9 * sE:
            assert v2 @ v1 / (np.linalg.norm(v2) * np.linalg.norm(v1)) == pytest.approx(
                cb / np.sqrt(r2b), abs=1e-10)
        assert np.allclose(vion, VTH_AR * box_muller(u[5:9]), rtol=1e-13, atol=1e-14 * VTH_AR)
        # where both directions have unit length the split is exact
        uq = u.copy()
        uq[2] = uq[4] = 0.25
        _, v1q, _, v2q = O.coll_kin(3, ME, MAR, M_EV, Eiz, VTH_AR, v, uq)
        e1q, e2q = 0.5 * M_EV * (v1q @ v1q), 0.5 * M_EV * (v2q @ v2q)
        assert e1q + e2q == pytest.approx(sE, rel=1e-12)
        assert e1q == pytest.approx(sE * u[0], rel=1e-11)
    assert seen > 100


def test_mcc_mirror_consumes_the_documented_draws(orc):
    """mir_mcc (bit-identical to the GPU kernels) applies mir_coll_kin to the
    uniforms of Philox blocks 1..5 of counter (id, step, stream, block):
    recompute each collided particle from its draws."""
    import ctypes as C
    from iskra_amd import problems as P
    xs = P.ccp_cross_sections()
    n = 3000
    rng = np.random.default_rng(40)
    types = [0, 2, 3]
    off, eps_all, sig_all = [0], [], []
    for e, s in xs["electron"]:
        eps_all += list(e)
        sig_all += list(s)
        off.append(len(eps_all))
    off = np.asarray(off, np.int32)
    eps_all, sig_all = np.asarray(eps_all), np.asarray(sig_all)
    a = O.MirMcc()
    a.ncoll = 3
    for q, t in enumerate(types):
        a.types[q] = t
        a.energy[q] = xs["energies"][q]
    a.off, a.eps, a.sig = O.ptr(off), O.ptr(eps_all), O.ptr(sig_all)
    a.candP, a.dt, a.m_eV, a.ms, a.mt = 0.9, 1e-11, M_EV, ME, MAR
    a.vth_t, a.tqm = VTH_AR, 0.0
    a.k0, a.k1, a.step, a.stream, a.dim, a.nx = 123, 456, 7, 0x4D430000, 1, 65
    cell = rng.integers(0, 64, n).astype(np.int32)
    ox = rng.random(n).astype(np.float32)
    oy = np.zeros(n, np.float32)
    v0 = (rng.standard_normal((3, n)) * 4e6).astype(np.float32)
    vx, vy, vz = [np.ascontiguousarray(q) for q in v0.copy()]
    ids = rng.permutation(10 * n)[:n].astype(np.uint32)
    E = np.zeros(65, np.float32)
    tn = np.full(65, 5e23)  # n sigma g dt <= 1.3, so P_k stays below 1
    nu = np.zeros(65 * 3, np.int64)
    ev = np.zeros(8, np.int64)
    pc, px, py = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(n, np.float32)
    pe, pi = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32)
    nprod = O.lib().mir_mcc(C.byref(a), n, O.ptr(cell), O.ptr(ox), O.ptr(oy), O.ptr(vx),
                            O.ptr(vy), O.ptr(vz), O.ptr(ids), O.ptr(E), O.ptr(tn), O.ptr(nu),
                            O.ptr(ev), O.ptr(pc), O.ptr(px), O.ptr(py), O.ptr(pe), O.ptr(pi),
                            None)
    assert nprod > 0 and ev[0] > 0 and ev[1] > 0
    changed = np.nonzero((vx != v0[0]) | (vy != v0[1]) | (vz != v0[2]))[0]
    assert len(changed) > 100
    g = 0
    for p in range(n):
        blk = [O.philox4x32([ids[p], 7, 0x4D430000, b], [123, 456]) for b in range(6)]
        u = np.array([O.u01(w[h], w[h + 1]) for w in blk for h in (0, 2)])
        if not u[0] < 0.9:
            assert p not in changed
            continue
        k = min(int(np.floor(3 * u[1])), 2)
        if p not in changed:
            continue
        vin = np.array([v0[0][p], v0[1][p], v0[2][p]], dtype=np.float64)
        used, v1, vion, v2 = O.coll_kin(types[k], ME, MAR, M_EV, xs["energies"][k], VTH_AR,
                                        vin, u[2:12])
        assert used > 0
        assert np.array_equal(np.array([vx[p], vy[p], vz[p]]), v1.astype(np.float32))
        if types[k] == 3:
            assert np.array_equal(pe[3 * g:3 * g + 3], v2.astype(np.float32))
            assert np.array_equal(pi[3 * g:3 * g + 3], vion.astype(np.float32))
            assert pc[g] == cell[p] and px[g] == ox[p]
            g += 1
    assert g == nprod
