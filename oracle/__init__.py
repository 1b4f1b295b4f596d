"""CPU oracle of the electrostatic PIC step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker; the product (iskra_amd/) never does.

The reference (bchaber/iskra) is Julia, and no Julia toolchain exists in this
environment, so the reference cannot be built or run: this oracle is a C
restatement (oracle/oracle.c) of the reference's algorithms, read from its
source, plus Python orchestration.

How each oracle function is pinned (checks live in tests/):

  ref_cell_volume        closed form (edge halving, corner quarter) and the
                         restated RegularGrids.jl rule; uniform plasma gives
                         uniform density including edges/corners.
  ref_particle_to_grid   single-particle closed forms; charge conservation
                         sum(n V) = sum(w); brute force on a hand table.
  ref_grid_to_particle   exact for linear fields (CIC interpolates linears).
  ref_push               closed form for B = 0 (v += qm E dt, x += v dt) and a
                         uniform-E trajectory; Boris rotation preserves |v|
                         when E = 0.
  ref_wrap / ref_discard closed forms of fld; swap-remove order on tiny arrays
                         enumerated by hand.
  ref_poisson_* + LU     FiniteDifferenceMethod/test/runtests.jl parallel-plate
                         capacitor (Ex = -1, Ey = 0), manufactured solutions,
                         LU residual ~ 1e-15, agreement with numpy.linalg.solve.
  ref_efield             closed form on linear / quadratic phi.
  ref_sigma              interpolation nodes, midpoints, flat extrapolation.
  ref_max_sigma_g        brute force over a dense energy grid (upper bound at
                         nodes, as the reference defines it).
  orc_philox4x32         Random123 known-answer vectors; structure cross-checked
                         through orc_philox4x64 against numpy.random.Philox.
  orc_exp/log/sincos2pi  libm within a few ulp over dense sweeps.
  orc_fix32/fixprod32    (the integer fixed-point CIC weights of the particle
                         kernels) against rint of the exact FP64 products on
                         edge cases and random offsets.
  mir_scatter            the reference's unrotated matrix literal applied as
                         written (row vector times matrix, numpy); chi = 0
                         returns the incident direction; |r|^2 and r.v_hat
                         against their closed forms for that (non-orthogonal)
                         matrix; unit length and r.v_hat = cos chi where cos
                         eta = 0; the chi distributions by KS tests against
                         their closed-form CDFs.
  mir_coll_kin           closed forms per collision type with the draws
                         injected: Box-Muller target velocity, centre-of-mass
                         velocity fixed, |v' - w| = mr2 |v_r| |r|, energy
                         conserved by elastic scattering where |r| = 1,
                         excitation eps' = (eps - E)|r|^2, ionisation energy
                         split e1 + e2 = eps - E, secondary relative to the
                         updated primary (tests/test_oracle_mcc_kinematics.py).
  mir_*                  against the ref_* functions within FP32 tolerance
                         (tests/test_oracle.py), and bit-exactly against the GPU.
  poisson_sparse         equal to the dense reference assembly on small grids;
                         its direct solve reproduces a linear potential.
  RefSim                 the reference's FP64 time loop composed of the ref_*
                         operators above; the FP32 twin stays within the stated
                         per-field bounds of it on the two-stream, slab and
                         discharge configurations, and a run with a planted
                         boundary error does not (tests/test_oracle_fp64_loop.py).
  MirSim (2-D)           the 1-D twin's operators on a 2-D grid; bit for bit
                         against the device over 40 steps with the device's
                         potential handed over, that potential within 1e-12 of
                         the FP64 banded solve (tests/test_gpu_loop.py).
  ref_cell_volume_axial  closed forms: the annulus volumes 2 pi r_i dr dz
                         (interior), pi (dr/2)^2 dz (axis), pi dz (r^2 -
                         (r - dr/2)^2) (outer wall); they sum to the cylinder
                         pi R^2 L (tests/test_oracle_rz.py).
  ref_poisson_assemble_rz manufactured solutions: the axisymmetric operator
                         is exact on phi = r^2 (Laplacian 4) and z^2 (2) on
                         interior rows, and row by row on the axis and wall
                         rows' mirrored forms; equal to poisson_sparse_rz;
                         W A symmetric with the stated row weights.
  ref_rz_transform       closed forms: r' = sqrt(x^2 + (dt v_theta)^2),
                         v_r'^2 + v_theta'^2 = v_r^2 + v_theta^2, a particle
                         with v_theta = 0 keeps its Cartesian step; the
                         axis-crossing sign quirk (cos >= 0) by hand.
  mir_push_deposit_rz    against ref_push + ref_rz_transform in FP32
                         tolerance, bit for bit against the device.
  mir_dsmc               closed forms with the draws it takes: floor(Nc) trials
                         per cell with the remainder carried (all accepted
                         when sigma(g) g >= sgmax), momentum and kinetic
                         energy of every cell conserved by each collision,
                         post-collision relative velocity |g| (B, A cos C,
                         A sin C) with B and C uniform (KS tests), a source
                         index equal to the target index redrawn only when
                         the species differ (tests/test_oracle_dsmc.py).
  surfaces.SurfaceTracker the reference's track!/check!/hit! restated on
                         absolute FP64 positions: closed forms of absorption,
                         mirror reflection and corner double reflection,
                         untracked cells left alone (tests/test_oracle_surfaces.py).
  mir_push_deposit_track against SurfaceTracker within FP32 tolerance, bit
                         for bit against the device.
  surfaces.electrode_system the reference's add_new_dof + apply_neumann rows:
                         a floating wall electrode over a grounded wall with
                         no charge gives the linear potential of slope -sigma
                         its rows state; the segment-end rows agree with it.
  surfaces.CircuitRLC    advance_circuit! converges to the closed-form RL
                         step response and RC discharge as dt -> 0.
  ref_advection_diffusion closed forms of the discrete operators: a sine
                         mode's Laplacian in 1-D; in 2-D a uniform density in
                         a linear drift (the reference divides both axes'
                         terms by 2 dx: dx != dy tells it from dy), a linear
                         density in a constant drift, x^2 + y^2 with the
                         reference's dx dy divisor; a literal loop
                         transcription bit for bit (tests/test_oracle_fluid.py).
  RefSim fluids          advance! of a fluid from the previous step's E (the
                         first step sees E = 0) against the restatement.
  openpmd_layout         (oracle/openpmd_layout.py) files written by the
                         library and read back with h5py compared object by
                         object; three planted layout mistakes caught; the
                         unit dimensions against the SI definitions of the
                         units (tests/test_openpmd.py).
  orc_ws_run             (learned warm start, oracle/ws_oracle.c; no reference
                         counterpart) FP64 gradients against central finite
                         differences; convolution, pooling and upsampling
                         against torch.nn.functional on the CPU (a library
                         routine, not the reference); the network is linear
                         in the output layer and positively homogeneous.

parity unpinned: none of the above can be checked against the reference's own
run (Julia absent).  The reference's MersenneTwister streams, LAPACK round-off
and Interpolations.jl round-off are not reproducible here; the MCC
candidate-selection rule is pinned statistically (expected event counts), not
bitwise, against the reference rule restated in ref_mcc_literal().
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            from . import build as _b  # noqa
            _b.build()
        _lib = C.CDLL(LIBPATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
I32, I64, U32, U64, D, F = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double, C.c_float


class MirMcc(C.Structure):
    _fields_ = [("ncoll", C.c_int), ("types", I32 * 8), ("energy", D * 8), ("off", P),
                ("eps", P), ("sig", P), ("candP", D), ("dt", D), ("m_eV", D), ("ms", D),
                ("mt", D), ("vth_t", D), ("tqm", D), ("k0", U32), ("k1", U32), ("step", U32),
                ("stream", U32), ("dim", C.c_int), ("nx", C.c_int)]


def _declare(L):
    sig = {
        "orc_philox4x32": (None, [P, P, P]),
        "orc_philox4x64": (None, [P, P, P]),
        "orc_u01": (D, [U32, U32]),
        "orc_exp": (D, [D]),
        "orc_log": (D, [D]),
        "orc_sincos2pi": (None, [D, P, P]),
        "ref_cell_volume": (None, [C.c_int, C.c_int, C.c_int, D, D, P, P]),
        "ref_particle_to_grid": (None, [C.c_int, I64, P, P, C.c_int, C.c_int, D, D, P]),
        "ref_grid_to_particle": (None, [C.c_int, I64, P, C.c_int, C.c_int, D, D, P, P]),
        "ref_push": (None, [I64, C.c_int, P, P, P, P, D, D]),
        "ref_wrap": (None, [I64, C.c_int, P, P, P, P, C.c_int, P]),
        "ref_discard": (I64, [P, I64, C.c_int, P, P, P, P, D, P, P, P, C.c_int, P]),
        "ref_poisson_assemble": (None, [C.c_int, C.c_int, D, P]),
        "ref_cell_volume_axial": (None, [C.c_int, C.c_int, D, D, P, P]),
        "ref_poisson_assemble_rz": (None, [C.c_int, C.c_int, D, D, P]),
        "ref_rz_transform": (None, [I64, P, P, D]),
        "mir_push_deposit_rz": (I64, [I64, P, P, P, P, P, P, P, C.c_int, C.c_int, C.c_int, F, D,
                                      F, C.c_int, C.c_int, P]),
        "mir_push_deposit_track": (I64, [I64, P, P, P, P, P, P, C.c_int, C.c_int, C.c_int, F, F, F,
                                         C.c_int, C.c_int, P, P, D, D, P, P]),
        "ref_apply_periodic_1d": (None, [C.c_int, D, P]),
        "ref_apply_periodic_2d": (None, [C.c_int, C.c_int, D, D, C.c_int, P]),
        "ref_apply_dirichlet": (None, [C.c_int, P, D, P, P, P]),
        "ref_lu_solve": (C.c_int, [C.c_int, P, P, P]),
        "ref_efield": (None, [C.c_int, C.c_int, C.c_int, D, D, P, P]),
        "ref_poisson_band_solve": (C.c_int, [C.c_int, C.c_int, D, P, P, P, D, P]),
        "ref_sigma": (D, [C.c_int, P, P, D]),
        "ref_max_sigma_g": (D, [C.c_int, P, P, P, D, C.c_int, P]),
        "mir_gather": (None, [C.c_int, I64, P, P, P, P, C.c_int, P, P]),
        "mir_push_deposit": (I64, [C.c_int, I64, P, P, P, P, P, P, C.c_int, C.c_int, C.c_int,
                                   F, F, F, C.c_int, C.c_int, C.c_int, P]),
        "mir_rho": (None, [C.c_int, C.c_int, P, P, P, P, P, P]),
        "mir_tri_factor": (C.c_int, [C.c_int, P, C.c_int, P, P, P]),
        "mir_tri_solve": (None, [C.c_int, P, P, P, P, P, P, D, C.c_int, P]),
        "mir_efield": (None, [C.c_int, C.c_int, C.c_int, D, D, P, P]),
        "mir_mcc": (I64, [C.POINTER(MirMcc), I64, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P,
                          P, P]),
        "ref_particles_omp": (I64, [I64, P, P, P, C.c_int, C.c_int, D, D, P, D, D, P,
                                    C.c_int]),
        "mir_scatter": (None, [C.c_int, D, D, D, D, D, P]),
        "mir_coll_kin": (C.c_int, [C.c_int, D, D, D, D, D, P, P, P]),
        "mir_sample": (None, [C.c_int, I64, I64, D, D, D, D, C.c_int, C.c_int, P, P, P, P, U64,
                              U32, P, P, P, P, P, P, P]),
        "mir_sort_tile": (I64, [C.c_int, I64, C.c_int, C.c_int, P, P, P, P, P, P, P]),
        "mir_sort_cell": (I64, [C.c_int, I64, C.c_int, C.c_int, P, P, P, P, P, P, P]),
        "mir_dsmc": (I64, [C.c_int, C.c_int, C.c_int, I64, P, P, P, P, I64, P, P, P, P, C.c_int,
                           D, D, D, D, D, D, D, D, C.c_int, P, P, P, P, U32, U32, U32, U32]),
        "orc_ws_param_count": (I64, [C.c_int, C.c_int]),
        "orc_fix32": (I64, [F]),
        "orc_fixprod32": (I64, [F, F]),
        "orc_ws_run": (D, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, P, P]),
        "orc_ws_run_obj": (D, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, D, D,
                               C.c_int, P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def ptr(a):
    """ctypes pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be contiguous"
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# RNG helpers
# --------------------------------------------------------------------------
def philox4x32(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32(ptr(c), ptr(k), ptr(o))
    return o


def philox4x64(ctr, key):
    c = np.asarray(ctr, dtype=np.uint64)
    k = np.asarray(key, dtype=np.uint64)
    o = np.zeros(4, dtype=np.uint64)
    lib().orc_philox4x64(ptr(c), ptr(k), ptr(o))
    return o


def u01(a, b):
    return lib().orc_u01(int(a), int(b))


# --------------------------------------------------------------------------
# reference-mode helpers (FP64, reference layouts)
# --------------------------------------------------------------------------
def cell_volume(dim, nx, ny, dx, dy, volper=(0, 0, 0, 0)):
    V = np.zeros(nx * ny)
    vp = np.asarray(volper, dtype=np.int32)
    lib().ref_cell_volume(dim, nx, ny, dx, dy, ptr(vp), ptr(V))
    return V


def cell_volume_axial(nr, nz, dr, dz, volper_z=(0, 0)):
    """cell_volume(::AxialGrid{2}) (RegularGrids.jl:40-53), r fastest."""
    V = np.zeros(nr * nz)
    vp = np.asarray(volper_z, dtype=np.int32)
    lib().ref_cell_volume_axial(nr, nz, dr, dz, ptr(vp), ptr(V))
    return V


def poisson_matrix_rz(nr, nz, dr, dz, dirichlet=()):
    """The reference's dense axisymmetric A (generalized_poisson.jl:70-199)
    with apply_dirichlet rows; returns A (column-major), b, isdof."""
    nn = nr * nz
    A = np.zeros(nn * nn)
    lib().ref_poisson_assemble_rz(nr, nz, dr, dz, ptr(A))
    b = np.zeros(nn)
    isdof = np.ones(nn, dtype=np.uint8)
    for mask, val in dirichlet:
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel(order="F")
        lib().ref_apply_dirichlet(nn, ptr(m), float(val), ptr(A), ptr(b), ptr(isdof))
    return A, b, isdof


def poisson_sparse_rz(nr, nz, dr, dz, tags, tagval):
    """The same reference system as poisson_matrix_rz for grids the dense
    assembly cannot hold (sparse, node k = i + nr j): radial second
    difference and +-0.5/(dr r) derivative terms on 0 < i < nr-1, axial second
    difference, a missing neighbour's link mirrored onto the diagonal;
    Dirichlet rows are the identity.  Returns (A, Dirichlet mask, values)."""
    import scipy.sparse as sps
    N = nr * nz
    k = np.arange(N)
    i, j = k % nr, k // nr
    t = np.asarray(tags, np.int8).ravel()
    free = t == 0
    cr, cz = 1.0 / dr ** 2, 1.0 / dz ** 2
    rows, cols, vals = [k[~free]], [k[~free]], [np.ones((~free).sum())]
    diag = np.where(free, -2.0 * cr - 2.0 * cz, 0.0)
    mid = (i > 0) & (i < nr - 1)
    rinv = np.where(mid, 0.5 / dr / np.maximum(i, 1) / dr, 0.0)
    for off, ok, c, extra in ((1, i < nr - 1, cr, rinv), (-1, i > 0, cr, -rinv),
                              (nr, j < nz - 1, cz, 0.0), (-nr, j > 0, cz, 0.0)):
        m = free & ok
        rows.append(k[m])
        cols.append(k[m] + off)
        vals.append((c + np.broadcast_to(extra, (N,)))[m])
        mirror = free & ~ok
        diag[mirror] += c
    rows.append(k[free])
    cols.append(k[free])
    vals.append(diag[free])
    A = sps.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(N, N))
    return A, ~free, np.asarray(tagval)[t]


def potential_rz(nr, nz, dr, dz, tags, tagval, rho, eps0):
    """calculate_electric_potential on the axisymmetric reference system,
    solved by sparse LU (FP64)."""
    import scipy.sparse.linalg as spla
    A, dmask, dval = poisson_sparse_rz(nr, nz, dr, dz, tags, tagval)
    b = np.where(dmask, dval, -np.asarray(rho, np.float64) / eps0)
    # the identity Dirichlet rows next to 1/dr^2 rows make a direct solve of
    # the full system lose ~1e-8 of max|phi|; the same solution from the free
    # rows with the Dirichlet columns moved to the right-hand side does not
    A = A.tocsr()
    free = ~dmask
    x = np.where(dmask, dval, 0.0).astype(np.float64)
    Aff = A[free][:, free].tocsc()
    rhs = b[free] - A[free][:, dmask] @ x[dmask]
    x[free] = spla.splu(Aff).solve(rhs)
    return x


def rz_transform(x, v, dt):
    """transform_from_cartesian_to_cylindrical! (pushers.jl:52-66) on
    reference-layout FP64 arrays (x: np x 2, v: np x 3, column-major), in place."""
    n = x.shape[0]
    xf = np.asfortranarray(x, dtype=np.float64)
    vf = np.asfortranarray(v, dtype=np.float64)
    lib().ref_rz_transform(n, xf.ctypes.data_as(C.c_void_p), vf.ctypes.data_as(C.c_void_p),
                           float(dt))
    x[...] = xf
    v[...] = vf


def particle_to_grid(dim, x, wg, nx, ny, dx, dy):
    x = np.ascontiguousarray(x, dtype=np.float64)
    np_ = len(wg)
    u = np.zeros(nx * ny)
    lib().ref_particle_to_grid(dim, np_, ptr(x), ptr(np.ascontiguousarray(wg, dtype=np.float64)),
                               nx, ny, dx, dy, ptr(u))
    return u


def grid_to_particle(dim, x, nx, ny, dx, dy, E):
    """x: flat column-major np x D; E: flat column-major nx x ny x 3."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    np_ = len(x) // dim
    pE = np.zeros(3 * np_)
    lib().ref_grid_to_particle(dim, np_, ptr(x), nx, ny, dx, dy,
                               ptr(np.ascontiguousarray(E, dtype=np.float64)), ptr(pE))
    return pE


def poisson_matrix(dim, nx, ny, dx, dy, periodic=(), dirichlet=()):
    """Assemble the reference's dense A and b (generalized_poisson.jl).
    periodic: list of axes as apply_periodic takes them (1-D: [1]);
    dirichlet: list of (mask nodes bool array, value).  Returns A, b, isdof."""
    nn = nx * ny
    A = np.zeros(nn * nn)
    lib().ref_poisson_assemble(nx, ny, dx, ptr(A))
    for ax in periodic:
        if dim == 1:
            lib().ref_apply_periodic_1d(nx, dx, ptr(A))
        else:
            lib().ref_apply_periodic_2d(nx, ny, dx, dy, ax, ptr(A))
    b = np.zeros(nn)
    isdof = np.ones(nn, dtype=np.uint8)
    for mask, val in dirichlet:
        m = np.ascontiguousarray(mask, dtype=np.uint8).ravel(order="F")
        lib().ref_apply_dirichlet(nn, ptr(m), float(val), ptr(A), ptr(b), ptr(isdof))
    return A, b, isdof


def poisson_sparse(nodes, dx, tags, tagval):
    """The reference system of create_poisson_solver + apply_dirichlet
    (generalized_poisson.jl:34-68,205-215) on a square grid, as a sparse
    matrix for grids the dense assembly cannot hold: natural 5-point rows
    scaled by 1/dx^2, Dirichlet rows replaced by the identity with b = the
    tag's value; node k = i + nodes*j.  Returns (A, Dirichlet mask, values)."""
    import scipy.sparse as sps
    N = nodes * nodes
    k = np.arange(N)
    i, j = k % nodes, k // nodes
    t = np.asarray(tags, np.int8).ravel()
    free = t == 0
    rows, cols, vals = [k[~free]], [k[~free]], [np.ones((~free).sum())]
    diag = np.zeros(N)
    for off, ok in ((1, i < nodes - 1), (-1, i > 0), (nodes, j < nodes - 1), (-nodes, j > 0)):
        m = free & ok
        rows.append(k[m])
        cols.append(k[m] + off)
        vals.append(np.full(m.sum(), 1.0 / dx ** 2))
        diag[m] -= 1.0 / dx ** 2
    rows.append(k[free])
    cols.append(k[free])
    vals.append(diag[free])
    A = sps.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                       shape=(N, N))
    return A, ~free, np.asarray(tagval)[t]


def lu_solve(A, b):
    n = len(b)
    x = np.zeros(n)
    sing = lib().ref_lu_solve(n, ptr(np.ascontiguousarray(A)), ptr(np.ascontiguousarray(b)), ptr(x))
    return x, sing


def potential(A, b, isdof, rho, eps0):
    """calculate_electric_potential(ps, -rho): b[dofs] = -rho/eps0, x = A\\b."""
    bb = b.copy()
    bb[isdof == 1] = -rho[isdof == 1] / eps0
    return lu_solve(A, bb)


def potential_banded(nx, ny, dx, tags, tagval, rho, eps0):
    """Direct solve of the reference system for a non-periodic grid of any size
    (band storage; the reference's dense A\\b is infeasible beyond ~4k nodes)."""
    t = np.ascontiguousarray(tags, dtype=np.int8).ravel()
    tv = np.zeros(17)
    tv[:len(tagval)] = tagval
    x = np.zeros(nx * ny)
    rc = lib().ref_poisson_band_solve(nx, ny, dx, ptr(t), ptr(tv),
                                      ptr(np.ascontiguousarray(rho, dtype=np.float64)), eps0,
                                      ptr(x))
    assert rc == 0, "banded solve failed"
    return x


def efield(dim, nx, ny, dx, dy, phi):
    E = np.zeros(nx * ny * 3)
    lib().ref_efield(dim, nx, ny, dx, dy, ptr(np.ascontiguousarray(phi, dtype=np.float64)), ptr(E))
    return E


def sigma(eps, sig, e):
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    sig = np.ascontiguousarray(sig, dtype=np.float64)
    return lib().ref_sigma(len(eps), ptr(eps), ptr(sig), float(e))


def max_sigma_g(tables, m_eV):
    off = [0]
    eps_all, sig_all = [], []
    for e, s in tables:
        eps_all += list(e)
        sig_all += list(s)
        off.append(len(eps_all))
    off = np.asarray(off, dtype=np.int32)
    eps_all = np.asarray(eps_all, dtype=np.float64)
    sig_all = np.asarray(sig_all, dtype=np.float64)
    work = np.zeros(len(eps_all) + 1)
    return lib().ref_max_sigma_g(len(tables), ptr(off), ptr(eps_all), ptr(sig_all), m_eV,
                                 len(work), ptr(work))


def ref_push(x, v, E, B, qm, dt, D):
    np_ = len(v) // 3
    lib().ref_push(np_, D, ptr(x), ptr(v), ptr(E), ptr(B), qm, dt)


def ref_mcc_literal(rng, np_, candP_per_particle_fn, N, max_P, remainder=0.0):
    """The reference's candidate rule (mcc.jl:248-251): Nc = floor(N max_P np +
    remainder) picks by index with replacement.  Returns (picks, remainder)."""
    val = N * max_P * np_ + remainder
    Nc = int(np.floor(val))
    rem = val - Nc
    picks = rng.integers(0, np_, size=Nc)
    return picks, rem


# --------------------------------------------------------------------------
# mirror mode: a CPU twin of the GPU step (1-D and 2-D operators; full 1-D loop)
# --------------------------------------------------------------------------
QE_MCC = 1.60217646e-19
KB_MCC = 1.3806503e-23


def scatter(kind, v, u1, u2):
    """mir_scatter: direction of a particle scattered from velocity v
    (kind 0 the reference's "isotropic" angles, 1 its diffuse reflection)."""
    r = np.zeros(3)
    lib().mir_scatter(int(kind), float(v[0]), float(v[1]), float(v[2]), float(u1), float(u2),
                      ptr(r))
    return r


def coll_kin(type_, ms, mt, m_eV, energy, vth_t, v, u):
    """mir_coll_kin: one collision (FP64) from velocity v and the uniforms u.
    Returns (uniforms used, velocity after, target/ion velocity, secondary)."""
    vv = np.ascontiguousarray(v, dtype=np.float64)
    uu = np.zeros(10)
    uu[:len(u)] = u
    out = np.zeros(9)
    k = lib().mir_coll_kin(int(type_), ms, mt, m_eV, energy, vth_t, ptr(vv), ptr(uu), ptr(out))
    return k, out[0:3].copy(), out[3:6].copy(), out[6:9].copy()


class MirSpecies:
    def __init__(self, dim, cap, q, m, w0, track_ids=True, bnd=(1, 1)):
        self.dim, self.cap, self.q, self.m, self.w0 = dim, cap, q, m, w0
        self.cell = np.zeros(cap, dtype=np.int32)
        self.ox = np.zeros(cap, dtype=np.float32)
        self.oy = np.zeros(cap, dtype=np.float32)
        self.vx = np.zeros(cap, dtype=np.float32)
        self.vy = np.zeros(cap, dtype=np.float32)
        self.vz = np.zeros(cap, dtype=np.float32)
        self.id = np.zeros(cap, dtype=np.uint32)
        self.np = 0
        self.next_id = 0
        self.track_ids = track_ids
        self.bnd = bnd
        self.cnt = None

    def set_state(self, cell, ox, oy, vx, vy, vz, ids):
        n = len(cell)
        self.cell[:n] = cell
        self.ox[:n] = ox
        if oy is not None:
            self.oy[:n] = oy
        self.vx[:n] = vx
        self.vy[:n] = vy
        self.vz[:n] = vz
        if ids is not None:
            self.id[:n] = ids
            self.next_id = int(ids.max()) + 1 if n else 0
        else:
            self.id[:n] = np.arange(n, dtype=np.uint32)
            self.next_id = n
        self.np = n

    def live(self):
        n = self.np
        m = self.cell[:n] >= 0
        return m


class MirSim:
    """CPU twin of iskra_step: same number format, same operation order, same
    particle order (stable tile sort every sort_every steps, per species when
    a species sets its own interval; products appended in parent order and
    numbered in the order of their parents' ids), same Philox draws.  In 1-D
    the Poisson solve is the library's direct solver, bit for bit.  In 2-D
    (ny > 1) the potential comes from the reference matrix solved directly in
    FP64 (banded LU: set_poisson_2d) -- or is set from outside between
    advance() and end_step() -- and E is formed from it as the device does."""

    def __init__(self, nx, dx, x0=0.0, volper=(0, 0, 0, 0), T=64, sort_every=4, seed=0,
                 ny=1, dy=1.0, axial=False):
        self.dim = 2 if ny > 1 else 1
        self.nx, self.ny = nx, ny
        self.ncx, self.ncy = nx - 1, (ny - 1 if self.dim == 2 else 1)
        self.dx, self.dy, self.x0 = dx, dy, x0
        self.axial = axial  # (r, z) grid with the RZ pusher
        if axial:
            self.vol = cell_volume_axial(nx, ny, dx, dy, volper[2:4])
        else:
            self.vol = cell_volume(self.dim, nx, ny, dx, dy, volper)
        self.T, self.sort_every = T, sort_every
        self.ntx = (self.ncx + T - 1) // T
        nn = nx * ny
        self.species = []
        self.bg = np.zeros(nn)
        self.rho = np.zeros(nn)
        self.phi = np.zeros(nn)
        self.E = np.zeros(nn * self.dim, dtype=np.float32)  # (Ex, Ey) per node in 2-D
        self.seed = seed
        self.step_count = 0
        self.mccs = []
        self.tags = np.zeros(nx, dtype=np.int8)
        self.tagval = np.zeros(17)
        self.wf = {}
        self.per_x = 0
        self.eps0 = 1.0
        self.pushed = 0

    def add_species(self, sp):
        self.species.append(sp)
        return len(self.species) - 1

    def set_poisson(self, eps0, tags, per_x):
        self.eps0 = eps0
        self.tags = np.ascontiguousarray(tags, dtype=np.int8)
        self.per_x = per_x
        self.singular = int(per_x and not np.any(self.tags))
        n = self.nx
        self.a, self.cp, self.inv = np.zeros(n), np.zeros(n), np.zeros(n)
        rc = lib().mir_tri_factor(n, ptr(self.tags), self.singular, ptr(self.a), ptr(self.cp),
                                  ptr(self.inv))
        assert rc == 0

    def set_poisson_2d(self, eps0, tags, tagval):
        """Dirichlet tags per node (x fastest) and tag values: the reference
        system (generalized_poisson.jl:34-68,205-215) solved by banded LU."""
        self.eps0 = eps0
        self.tags = np.ascontiguousarray(tags, dtype=np.int8).ravel()
        self.tagval = np.zeros(17)
        self.tagval[:len(tagval)] = tagval

    # -- surface tracker, electrodes, circuit (the device's arithmetic) --
    def set_tracker(self, surf, tracked):
        """Edge codes (4 per cell: -x, +x, -y, +y) and tracked cells as the
        device holds them (0 none, 1 absorb, 2 reflect, 3 pass, 16 + k
        floating electrode k)."""
        self.surf = np.ascontiguousarray(surf, dtype=np.uint8)
        self.tracked = np.ascontiguousarray(tracked, dtype=np.uint8)

    def set_electrodes(self, sigma, area, circuit=None):
        """Floating electrodes' initial sigma and area (per sigma dof); circuit:
        dict(R, L, C, v0, v1, f, ph, pos_ref, pos_dof, neg_phi) or None."""
        self.el_sigma = np.zeros(8)
        self.el_dq = np.zeros(8)
        self.el_area = np.zeros(8)
        self.el_sigma[:len(sigma)] = sigma
        self.el_area[:len(area)] = area
        self.nfloat = len(sigma)
        self.circuit = circuit
        self.circ = np.zeros(4)  # i, q, t, vext

    def electrode_update(self, dt, with_circuit):
        """The device's k_el_update: hits -> charge, then the circuit."""
        for k in range(self.nfloat):
            dq = 0.0
            for sp in self.species:
                eh = getattr(sp, "ehit", None)
                if eh is not None and eh[k]:
                    dq += float(eh[k]) * (sp.q * sp.w0)
                    eh[k] = 0
            if dq != 0.0:
                self.el_dq[k] += dq
                self.el_sigma[k] += dq / self.el_area[k]
        c = self.circuit
        if not with_circuit or c is None:
            return
        i, q, t = self.circ[0], self.circ[1], self.circ[2]
        vext = (self.phi[c["pos_ref"]] - c["neg_phi"]) if c["pos_ref"] >= 0 else 0.0
        u = c["f"] * t + c["ph"] / 6.283185307179586
        u = u - np.floor(u)
        sn, cn = C.c_double(), C.c_double()
        lib().orc_sincos2pi(u, C.byref(sn), C.byref(cn))
        vt = c["v0"] + c["v1"] * sn.value
        ni = (c["L"] / dt - c["R"] / 2.0) * i + vext - vt
        nq = q
        if c["C"] > 0.0:
            ni -= q / c["C"]
            ni /= (c["L"] / dt + c["R"] / 2.0)
            nq = q + dt * i
        else:
            ni /= (c["L"] / dt + c["R"] / 2.0)
        self.circ[:] = (ni, nq, t + dt, vext)
        if c["pos_dof"] >= 0:
            self.el_dq[c["pos_dof"]] = 0.0
            self.el_sigma[0] += -dt * ni / self.el_area[c["pos_dof"]]

    def add_dsmc(self, src, tgt, tab_g, tab_sigma):
        """One DSMC.ElasticCollision interaction (equal weights), run after
        the MCC interactions of each step, as iskra_step orders them."""
        g = np.ascontiguousarray(tab_g, dtype=np.float64)
        sg = np.ascontiguousarray(tab_sigma, dtype=np.float64)
        km = int(np.argmax(sg))
        d = dict(src=src, tgt=tgt, g=g, sig=sg, sgmax=float(sg[km] * g[km]),
                 rem=np.zeros(self.ncx * self.ncy), nu=np.zeros(self.nx * self.ny, np.int64),
                 step=0, stream=0x44530000 + len(getattr(self, "dsmcs", [])), last=0)
        if not hasattr(self, "dsmcs"):
            self.dsmcs = []
        self.dsmcs.append(d)
        return len(self.dsmcs) - 1

    def dsmc_run(self, k, dt):
        d = self.dsmcs[k]
        A, B = self.species[d["src"]], self.species[d["tgt"]]
        same = d["src"] == d["tgt"]
        self.sort(d["src"], exact=True)
        if not same:
            self.sort(d["tgt"], exact=True)
        k0, k1 = self.seed & 0xFFFFFFFF, (self.seed >> 32) & 0xFFFFFFFF
        ev = lib().mir_dsmc(self.ncx, self.ncy, self.nx, A.np, ptr(A.cell), ptr(A.vx), ptr(A.vy),
                            ptr(A.vz), B.np, ptr(B.cell), ptr(B.vx), ptr(B.vy), ptr(B.vz),
                            int(same), A.w0, B.w0, self.dx, self.dy, dt, d["sgmax"],
                            A.m / (A.m + B.m), B.m / (A.m + B.m), len(d["g"]), ptr(d["g"]),
                            ptr(d["sig"]), ptr(d["rem"]), ptr(d["nu"]), d["step"], d["stream"],
                            k0, k1)
        d["step"] += 1
        d["last"] = ev
        return ev

    def add_mcc(self, src, ion, types, energy, tables, target_m, target_q, target_T, tn):
        off = [0]
        eps_all, sig_all = [], []
        for e, s in tables:
            eps_all += list(e)
            sig_all += list(s)
            off.append(len(eps_all))
        m = dict(src=src, ion=ion, types=list(types), energy=list(energy),
                 off=np.asarray(off, dtype=np.int32), eps=np.asarray(eps_all, dtype=np.float64),
                 sig=np.asarray(sig_all, dtype=np.float64), target_m=target_m,
                 target_q=target_q, target_T=target_T,
                 tn=np.ascontiguousarray(tn, dtype=np.float64), step=0,
                 stream=0x4D430000 + len(self.mccs))
        sp = self.species[src]
        m["m_eV"] = sp.m / QE_MCC
        m["max_sg"] = max_sigma_g(tables, m["m_eV"])
        m["nu"] = np.zeros(self.nx * self.ny * len(types), dtype=np.int64)
        m["events"] = np.zeros(8, dtype=np.int64)
        self.mccs.append(m)
        return len(self.mccs) - 1

    def mcc_run(self, k, dt):
        m = self.mccs[k]
        sp = self.species[m["src"]]
        N = len(m["types"])
        maxP = 1.0 - lib().orc_exp(-float(m["tn"].max()) * m["max_sg"] * dt)
        a = MirMcc()
        a.ncoll = N
        for q in range(N):
            a.types[q] = m["types"][q]
            a.energy[q] = m["energy"][q]
        a.off, a.eps, a.sig = ptr(m["off"]), ptr(m["eps"]), ptr(m["sig"])
        a.candP = N * maxP
        a.dt = dt
        a.m_eV = m["m_eV"]
        a.ms = sp.m
        a.mt = m["target_m"]
        a.vth_t = np.sqrt(2.0 * KB_MCC * m["target_T"] / m["target_m"])
        a.tqm = m["target_q"] / m["target_m"]
        a.k0 = self.seed & 0xFFFFFFFF
        a.k1 = (self.seed >> 32) & 0xFFFFFFFF
        a.step = m["step"]
        a.stream = m["stream"]
        a.dim = self.dim
        a.nx = self.nx
        n = sp.np
        pcell = np.zeros(max(n, 1), dtype=np.int32)
        pox = np.zeros(max(n, 1), dtype=np.float32)
        poy = np.zeros(max(n, 1), dtype=np.float32)
        pe = np.zeros(3 * max(n, 1), dtype=np.float32)
        pi = np.zeros(3 * max(n, 1), dtype=np.float32)
        ppid = np.zeros(max(n, 1), dtype=np.uint32)
        ev = np.zeros(8, dtype=np.int64)
        nprod = lib().mir_mcc(C.byref(a), n, ptr(sp.cell), ptr(sp.ox), ptr(sp.oy), ptr(sp.vx),
                              ptr(sp.vy), ptr(sp.vz), ptr(sp.id), ptr(self.E), ptr(m["tn"]),
                              ptr(m["nu"]), ptr(ev), ptr(pcell), ptr(pox), ptr(poy), ptr(pe),
                              ptr(pi), ptr(ppid))
        if nprod < 0:
            raise RuntimeError("mirror mcc: probability above 1")
        m["events"] += ev
        m["last_events"] = ev[:N].copy()
        m["step"] += 1
        if nprod:
            R = 0
            if m["ion"] is not None and m["ion"] >= 0:
                R = int(np.rint(sp.w0 / self.species[m["ion"]].w0))
            # products are stored in particle order but numbered in the order
            # of their parents' ids (the rank of the parent id)
            rank = np.empty(nprod, dtype=np.int64)
            rank[np.argsort(ppid[:nprod], kind="stable")] = np.arange(nprod)
            # electrons appended to the source species
            s0 = sp.np
            k = min(nprod, sp.cap - s0)
            sp.cell[s0:s0 + k] = pcell[:k]
            sp.ox[s0:s0 + k] = pox[:k]
            sp.oy[s0:s0 + k] = poy[:k]
            sp.vx[s0:s0 + k] = pe[0:3 * k:3]
            sp.vy[s0:s0 + k] = pe[1:3 * k:3]
            sp.vz[s0:s0 + k] = pe[2:3 * k:3]
            sp.id[s0:s0 + k] = (sp.next_id + rank[:k]).astype(np.uint32)
            sp.np += k
            sp.next_id += nprod
            if R > 0:
                ion = self.species[m["ion"]]
                i0 = ion.np
                g = np.repeat(np.arange(nprod), R)
                kk = min(len(g), ion.cap - i0)
                g = g[:kk]
                ion.cell[i0:i0 + kk] = pcell[g]
                ion.ox[i0:i0 + kk] = pox[g]
                ion.oy[i0:i0 + kk] = poy[g]
                ion.vx[i0:i0 + kk] = pi[3 * g]
                ion.vy[i0:i0 + kk] = pi[3 * g + 1]
                ion.vz[i0:i0 + kk] = pi[3 * g + 2]
                r_of = np.tile(np.arange(R), nprod)[:kk]
                ion.id[i0:i0 + kk] = (ion.next_id + rank[g] * R + r_of).astype(np.uint32)
                ion.np += kk
                ion.next_id += nprod * R
        return m["last_events"]

    def push(self, s, dt, do_push=1):
        sp = self.species[s]
        cnt = np.zeros(self.nx * self.ny, dtype=np.int64)
        qmdt = np.float32(sp.q / sp.m * dt)
        if getattr(self, "surf", None) is not None and do_push:
            if getattr(sp, "ehit", None) is None:
                sp.ehit = np.zeros(8, dtype=np.uint64)
            live = lib().mir_push_deposit_track(
                sp.np, ptr(sp.cell), ptr(sp.ox), ptr(sp.oy), ptr(sp.vx), ptr(sp.vy), ptr(self.E),
                self.nx, self.ncx, self.ncy, qmdt, np.float32(dt / self.dx),
                np.float32(dt / self.dy), sp.bnd[0], sp.bnd[1], ptr(self.surf), ptr(self.tracked),
                dt, self.dx, ptr(sp.ehit), ptr(cnt))
            self.pushed += live
            sp.cnt = cnt
            return
        if self.axial and do_push:
            live = lib().mir_push_deposit_rz(sp.np, ptr(sp.cell), ptr(sp.ox), ptr(sp.oy),
                                             ptr(sp.vx), ptr(sp.vy), ptr(sp.vz), ptr(self.E),
                                             self.nx, self.ncx, self.ncy, qmdt, dt / self.dx,
                                             np.float32(dt / self.dy), sp.bnd[0], sp.bnd[1],
                                             ptr(cnt))
            self.pushed += live
            sp.cnt = cnt
            return
        live = lib().mir_push_deposit(self.dim, sp.np, ptr(sp.cell), ptr(sp.ox), ptr(sp.oy),
                                      ptr(sp.vx), ptr(sp.vy), ptr(self.E), self.nx, self.ncx,
                                      self.ncy, qmdt, np.float32(dt / self.dx),
                                      np.float32(dt / self.dy), sp.bnd[0], sp.bnd[1], do_push,
                                      ptr(cnt))
        if do_push:
            self.pushed += live
        sp.cnt = cnt

    def compute_rho(self):
        ns = len(self.species)
        cnt = np.ascontiguousarray(np.stack([sp.cnt for sp in self.species]))
        w0 = np.array([sp.w0 for sp in self.species])
        q = np.array([sp.q for sp in self.species])
        lib().mir_rho(self.nx * self.ny, ns, ptr(cnt), ptr(w0), ptr(q), ptr(self.vol),
                      ptr(self.bg), ptr(self.rho))

    def solve(self):
        if self.axial:
            self.set_phi(potential_rz(self.nx, self.ny, self.dx, self.dy, self.tags, self.tagval,
                                      self.rho, self.eps0))
            return
        if self.dim == 2:
            self.set_phi(potential_banded(self.nx, self.ny, self.dx, self.tags, self.tagval,
                                          self.rho, self.eps0))
            return
        lib().mir_tri_solve(self.nx, ptr(self.rho), ptr(self.tags), ptr(self.tagval), ptr(self.a),
                            ptr(self.cp), ptr(self.inv), (self.dx * self.dx) / self.eps0,
                            self.singular, ptr(self.phi))
        lib().mir_efield(1, self.nx, 1, self.dx, 1.0, ptr(self.phi), ptr(self.E))

    def set_phi(self, phi):
        """Take phi (FP64, x fastest) and form E from it as the device does."""
        self.phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
        lib().mir_efield(self.dim, self.nx, self.ny, self.dx, self.dy, ptr(self.phi), ptr(self.E))

    def sort(self, s, exact=False):
        sp = self.species[s]
        f = lib().mir_sort_cell if exact else lib().mir_sort_tile
        sp.np = int(f(self.dim, sp.np, self.T, self.ntx, ptr(sp.cell), ptr(sp.ox), ptr(sp.oy),
                      ptr(sp.vx), ptr(sp.vy), ptr(sp.vz), ptr(sp.id)))

    def waveform(self, t):
        for tag, (off, amp, f, ph_turns) in self.wf.items():
            u = f * t + ph_turns
            u = u - np.floor(u)
            s = C.c_double()
            c = C.c_double()
            lib().orc_sincos2pi(u, C.byref(s), C.byref(c))
            self.tagval[tag] = off + amp * s.value

    def advance(self, dt):
        """The part of one step before the Poisson solve: MCC, push and
        deposit of every species, rho."""
        for k in range(len(self.mccs)):
            self.mcc_run(k, dt)
        for k in range(len(getattr(self, "dsmcs", []))):
            self.dsmc_run(k, dt)
        for s in range(len(self.species)):
            self.push(s, dt)
        if getattr(self, "el_area", None) is not None:
            self.electrode_update(dt, True)
        self.compute_rho()

    def end_step(self, dt):
        """The part after the solve: electrode waveforms, step count, sorts."""
        self.waveform(self.step_count * dt)
        self.step_count += 1
        for s, sp in enumerate(self.species):
            every = getattr(sp, "sort_every", 0) or self.sort_every
            if every > 0 and self.step_count % every == 0:
                self.sort(s)

    def step(self, dt, nsteps=1):
        for _ in range(nsteps):
            self.advance(dt)
            self.solve()
            self.end_step(dt)


def ref_advection_diffusion(n, D, v, dh, dt):
    """calculate_advection_diffusion(n, D, v, dh, dt)
    (FiniteDifferenceMethod/src/advection_diffusion.jl:4-38) in FP64, each
    expression evaluated as the reference writes it (left to right, 2dx as
    2*dx, dx^2 as dx*dx).  n: (nx,) or (nx, ny) indexed [i, j] like the
    reference's arrays; v: (nx, 3) or (nx, ny, 3); returns dn of n's shape,
    zero on the boundary nodes.  1-D: diffusion only (the advection lines
    are commented out there, :13-15)."""
    n = np.asarray(n, dtype=np.float64)
    d = np.zeros_like(n)
    if n.ndim == 1:
        dx = dh[0]
        if len(n) >= 3:
            d[1:-1] = 0.0 + ((dt * D) * ((n[:-2] - 2.0 * n[1:-1]) + n[2:])) / (dx * dx)
        return d
    dx, dy = dh[0], dh[1]
    if n.shape[0] < 3 or n.shape[1] < 3:
        return d
    vx, vy = v[..., 0], v[..., 1]
    c = (slice(1, -1), slice(1, -1))
    L, R = (slice(None, -2), slice(1, -1)), (slice(2, None), slice(1, -1))
    Dn, Up = (slice(1, -1), slice(None, -2)), (slice(1, -1), slice(2, None))
    lap = (((n[L] + n[Dn]) - 4.0 * n[c]) + n[R]) + n[Up]
    t = 0.0 + ((dt * D) * lap) / (dx * dy)
    t = t - ((dt * n[c]) * (((vx[R] - vx[L]) + vy[Up]) - vy[Dn])) / (2.0 * dx)
    t = t - (dt * ((vx[c] * (n[R] - n[L])) + (vy[c] * (n[Up] - n[Dn])))) / dx
    d[c] = t
    return d


class RefSim:
    """The reference's time loop in FP64 (ParticleInCell.jl:84-139) built from
    the ref_* restatements, with the reference's data model: absolute FP64
    positions x (D x np), FP64 velocities, weights wg.  One step: for every
    kinetic species grid_to_particle of E (cloud_in_cell.jl:20-36), the Boris
    push with B = 0 (pushers.jl:37-50), the after_push boundary (wrap!/discard!
    with swap-remove, wrap.jl:1-33, kinetic.jl:20-27); then density =
    particle_to_grid(wg) / cell_volume (kinetic.jl:53), rho = sum q n + the
    fluid background (ParticleInCell.jl:118-124), calculate_electric_potential
    (generalized_poisson.jl:372-378) as a direct solve of the reference matrix,
    calculate_electric_field (:385-410) and the electrode waveform of the
    after_loop hook at t = (it-1) dt (problem/11_rf_discharge.jl:95).

    Poisson: set_ring_1d() for the 1-D periodic ring (A1; the singular system's
    minimum-norm solution, i.e. mean-free phi, A2), set_tags() for Dirichlet
    tags solved by banded LU (any grid size).  No MCC (the reference's
    collisions draw from Julia's MersenneTwister, not reproducible here)."""

    def __init__(self, dim, nx, dx, ny=1, dy=1.0, eps0=1.0, volper=(0, 0, 0, 0)):
        self.dim, self.nx, self.ny, self.dx, self.dy, self.eps0 = dim, nx, ny, dx, dy, eps0
        self.vol = cell_volume(dim, nx, ny, dx, dy, volper)
        nn = nx * ny
        self.bg = np.zeros(nn)
        self.rho = np.zeros(nn)
        self.phi = np.zeros(nn)
        self.E = np.zeros(3 * nn)  # component-major, as ref_efield writes it
        self.species = []
        self.fluids = []
        self.pinv = None
        self.tags = None
        self.tagval = np.zeros(17)
        self.wf = {}
        self.step_count = 0

    def set_ring_1d(self):
        A, _, _ = poisson_matrix(1, self.nx, 1, self.dx, 1.0, periodic=[1])
        self.pinv = np.linalg.pinv(A.reshape(self.nx, self.nx, order="F"))

    def set_tags(self, tags, tagval=()):
        self.tags = np.ascontiguousarray(tags, dtype=np.int8).ravel()
        self.tagval[:] = 0.0
        self.tagval[:len(tagval)] = tagval

    def add_species(self, q, m, w0, x, v, bnd):
        """x: (np, D) FP64 positions, v: (np, 3); bnd per dimension (0 none,
        1 wrap, 2 discard)."""
        x = np.asarray(x, dtype=np.float64).reshape(len(v), -1)
        n = len(v)
        self.species.append(dict(
            q=q, m=m, w0=w0, np=n, bnd=tuple(bnd),
            x=np.ascontiguousarray(x.T), v=np.ascontiguousarray(np.asarray(v, np.float64).T),
            wg=np.full(n, float(w0)), id=np.arange(1, n + 1, dtype=np.uint32)))
        return len(self.species) - 1

    def add_fluid(self, q, m, mu, n):
        """A FluidSpecies (pic/fluid.jl:1-11) with node densities n (nodes,
        x fastest), advanced every step by advance!(::FluidSpecies)
        (ParticleInCell.jl:74-82)."""
        self.fluids.append(dict(q=q, m=m, mu=mu, n=np.array(n, dtype=np.float64).ravel(),
                                dn=np.zeros(self.nx * self.ny)))
        return len(self.fluids) - 1

    def _advance_fluid(self, fl, dt):
        """v = E dt (q/m) with the E the step starts from, n += dn"""
        nn, sh = self.nx * self.ny, (self.ny, self.nx)
        E = self.E.reshape(3, nn)
        v = np.stack([(E[k].reshape(sh).T * dt) * (fl["q"] / fl["m"]) for k in range(3)], axis=-1)
        n = fl["n"].reshape(sh).T
        if self.dim == 1:
            n, v = n[:, 0], v[:, 0, :]
        dn = ref_advection_diffusion(n, fl["mu"], v, (self.dx, self.dy), dt)
        fl["dn"] = np.ascontiguousarray(dn.reshape(self.nx, -1).T).ravel()
        fl["n"] = fl["n"] + fl["dn"]

    def _advance_species(self, sp, dt):
        D, n = self.dim, sp["np"]
        if n == 0:
            return
        x = np.ascontiguousarray(sp["x"][:, :n])
        v = np.ascontiguousarray(sp["v"][:, :n])
        pE = grid_to_particle(D, x.ravel(), self.nx, self.ny, self.dx, self.dy, self.E)
        B = np.zeros(3 * n)
        xf, vf = x.ravel().copy(), v.ravel().copy()
        ref_push(xf, vf, pE, B, sp["q"] / sp["m"], dt, D)
        x, v = xf.reshape(D, n), vf.reshape(3, n)
        nodes = np.array([self.nx, self.ny], dtype=np.int64)
        dh = np.array([self.dx, self.dy])
        org = np.zeros(2)
        wrapd = np.array([d for d in range(D) if sp["bnd"][d] == 1], dtype=np.int32)
        if len(wrapd):
            xf = np.ascontiguousarray(x).ravel()
            lib().ref_wrap(n, D, ptr(xf), ptr(nodes), ptr(dh), ptr(org), len(wrapd), ptr(wrapd))
            x = xf.reshape(D, n)
        disd = np.array([d for d in range(D) if sp["bnd"][d] == 2], dtype=np.int32)
        if len(disd):
            xf = np.ascontiguousarray(x).ravel()
            vf = np.ascontiguousarray(v).ravel()
            wg = np.ascontiguousarray(sp["wg"][:n])
            ids = np.ascontiguousarray(sp["id"][:n])
            npr = C.c_int64(n)
            lib().ref_discard(C.byref(npr), n, D, ptr(xf), ptr(vf), ptr(wg), ptr(ids), sp["w0"],
                              ptr(nodes), ptr(dh), ptr(org), len(disd), ptr(disd))
            n = npr.value
            x, v = xf.reshape(D, -1)[:, :n], vf.reshape(3, -1)[:, :n]
            sp["wg"], sp["id"] = wg[:n].copy(), ids[:n].copy()
        sp["x"], sp["v"], sp["np"] = np.ascontiguousarray(x), np.ascontiguousarray(v), n

    def density(self, s):
        sp = self.species[s]
        n = sp["np"]
        x = np.ascontiguousarray(sp["x"][:, :n]).ravel()
        return particle_to_grid(self.dim, x, sp["wg"][:n], self.nx, self.ny, self.dx,
                                self.dy) / self.vol

    def step(self, dt, nsteps=1):
        for _ in range(nsteps):
            for sp in self.species:
                self._advance_species(sp, dt)
            for fl in self.fluids:
                self._advance_fluid(fl, dt)
            rho = np.zeros(self.nx * self.ny)
            for s, sp in enumerate(self.species):
                rho += sp["q"] * self.density(s)
            for fl in self.fluids:
                rho += fl["q"] * fl["n"]
            self.rho = rho + self.bg
            if self.pinv is not None:
                self.phi = self.pinv @ (-self.rho / self.eps0)
            else:
                self.phi = potential_banded(self.nx, self.ny, self.dx, self.tags, self.tagval,
                                            self.rho, self.eps0)
            self.E = efield(self.dim, self.nx, self.ny, self.dx, self.dy, self.phi)
            t = self.step_count * dt
            for tag, (off, amp, f, ph) in self.wf.items():
                self.tagval[tag] = off + amp * np.sin(2.0 * np.pi * (f * t + ph))
            self.step_count += 1

    def moments(self, s):
        """np, sum x (per dimension), sum v, sum |v|^2 of species s."""
        sp = self.species[s]
        n = sp["np"]
        return dict(np=n, x=sp["x"][:, :n].sum(axis=1), v=sp["v"][:, :n].sum(axis=1),
                    v2=float((sp["v"][:, :n] ** 2).sum()))


# --------------------------------------------------------------------------
# learned warm start network (oracle/ws_oracle.c)
# --------------------------------------------------------------------------
def ws_param_count(L, base):
    return int(lib().orc_ws_param_count(L, base))


def ws_run(L, base, params, inputs, target=None, round_bf16=True, grads=True, objective=0,
           gain=1.0, omega=0.0, sweeps=1):
    """inputs: (B, S, S); returns (out (B,S,S), loss, grads or None).
    objective 0: mean squared error against target; 1: the energy objective
    (network times gain, then `sweeps` damped Jacobi sweeps of weight omega,
    scored in the A-norm relative to gain * target)."""
    x = np.ascontiguousarray(inputs, dtype=np.float64)
    B, S = x.shape[0], x.shape[1]
    p = np.ascontiguousarray(params, dtype=np.float64)
    assert p.size == ws_param_count(L, base)
    out = np.zeros_like(x)
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float64)
    g = np.zeros_like(p) if (grads and t is not None) else None
    loss = lib().orc_ws_run_obj(L, base, S, B, int(round_bf16), int(objective), float(gain),
                                0.25 * float(omega), int(sweeps), ptr(p), ptr(x), ptr(t),
                                ptr(out), ptr(g))
    return out, loss, g
