"""ParticleInCell: species, sources, the particle operators and the time loop
(ParticleInCell/src/ParticleInCell.jl, pic/*.jl) plus the set-up helpers of
problem/configuration.jl.

Species are host objects with the reference's fields (x: N x D, v: N x 3,
np, wg, id, q, m, w0, name).  Once a simulation binds them (`solve`, or
`bind`), their particles live on the GPU; the host arrays become a mirror
that is downloaded when read and uploaded again before the next device
operation if it was touched, so hooks may edit `part.x`, `part.v` and
`part.np` as reference scripts do.  Every particle operation runs on the GPU;
there is no host implementation of any of them.

Hooks are module attributes that a script replaces, as a Julia script
redefines ParticleInCell.after_push & co. (ParticleInCell.jl:38-41):

    import iskra_amd.particle_in_cell as pic
    pic.after_push = lambda part, grid: pic.wrap(part, grid)

`after_push` is evaluated once per species when the simulation is bound: the
wrap/discard calls it makes select the boundary that the fused push kernel
applies after every push (wrap.jl:1-33).
"""
import math

import numpy as np

from . import _lib as L
from . import diagnostics
from .finite_difference_method import PoissonSolver
from .regular_grids import UniformGrid

KB = 1.380649e-23  # problem/units_and_constants.jl (CODATA 2018, Unitful)


# ---------------------------------------------------------------------------
# species and sources (pic/kinetic.jl, pic/fluid.jl, pic/sources.jl)

class KineticSpecies:
    """KineticSpecies{D,V} (pic/kinetic.jl:1-18) with a device copy."""

    def __init__(self, name, N, D=2, V=3):
        if V != 3:
            raise ValueError("KineticSpecies: only V = 3 velocity components")
        self.name = name
        self.N = int(N)
        self.D = int(D)
        self.V = 3
        self.m = 0.0
        self.q = 0.0
        self.w0 = 1.0
        self.wg = np.ones(self.N)
        self._x = np.zeros((self.N, self.D))
        self._v = np.zeros((self.N, 3))
        self._id = np.arange(1, self.N + 1, dtype=np.uint32)  # particle_uuids (kinetic.jl:15)
        self._np = 0
        self.n = None            # last density (set by the loop)
        self._ctx = None
        self._sid = -1
        self._host_stale = False  # device holds newer particles than the host arrays
        self._dev_stale = False   # host arrays were handed out / changed since upload

    def __repr__(self):
        return self.name

    # host mirror -------------------------------------------------------------
    def _pull(self):
        if self._ctx is not None and self._host_stale:
            X, Vv, ids = self._ctx.download(self._sid, self.N)
            n = X.shape[0]
            self._x[:n] = X
            self._v[:n] = Vv
            self._id[:n] = ids
            self._np = n
            self.wg[:] = self.w0
            self._host_stale = False

    def _push(self):
        """Upload the host arrays if they may have changed."""
        if self._ctx is not None and self._dev_stale:
            n = self._np
            self._ctx.upload(self._sid, self._x[:n], self._v[:n], self._id[:n])
            self._dev_stale = False

    def _touched(self):
        self._pull()
        if self._ctx is not None:
            self._dev_stale = True

    @property
    def x(self):
        self._touched()
        return self._x

    @x.setter
    def x(self, val):
        self._touched()
        self._x[...] = val

    @property
    def v(self):
        self._touched()
        return self._v

    @v.setter
    def v(self, val):
        self._touched()
        self._v[...] = val

    @property
    def id(self):
        self._touched()
        return self._id

    @property
    def np(self):
        if self._ctx is not None and self._host_stale:
            return self._ctx.count(self._sid)
        return self._np

    @np.setter
    def np(self, val):
        self._touched()
        if not 0 <= int(val) <= self.N:
            raise ValueError(f"{self.name}: np = {val} outside [0, {self.N}]")
        self._np = int(val)


class FluidSpecies:
    """FluidSpecies{D} (pic/fluid.jl:1-11): density n (per node), mobility
    mu, charge q, mass m, temperature T.  It enters rho as q*n (no volume
    division, fluid.jl:11) and may be the MCC target gas.  A fluid that the
    reference's update cannot change (mu = 0, and in 2-D also q = 0:
    advection_diffusion.jl:4-38) is a constant background on the device;
    any other is advanced there every step (iskra_fluid_add), and n is then
    a host mirror of the device density, downloaded when read and uploaded
    before the next device call when changed, as the kinetic species'
    arrays are."""

    def __init__(self, name, mu, q, m, n, T):
        self.name = name
        self.mu = float(mu)
        self.q = float(q)
        self.m = float(m)
        self._n = np.array(n, dtype=np.float64)
        self.T = float(T)
        self._ctx = None
        self._fid = -1
        self._grid_n = None
        self._host_stale = False
        self._dev_stale = False

    def __repr__(self):
        return self.name

    def _pull(self):
        if self._ctx is not None and self._host_stale:
            self._n[...] = self._ctx.field(L.F_FLUID_N, self._fid).reshape(
                self._grid_n[::-1]).T
            self._host_stale = False

    def _push(self):
        if self._ctx is not None and self._dev_stale:
            self._ctx.set_field(L.F_FLUID_N, np.ascontiguousarray(self._n.T).ravel(),
                                index=self._fid)
            self._dev_stale = False

    @property
    def n(self):
        if self._ctx is not None:
            self._pull()
            self._dev_stale = True
        return self._n

    @n.setter
    def n(self, val):
        if self._ctx is not None:
            self._pull()
            self._dev_stale = True
            self._n[...] = val
        else:
            self._n = np.array(val, dtype=np.float64)


def create_fluid_species(name, mu, q, m, mx, my=None, T=300.0):
    """create_fluid_species (problem/configuration.jl:104-108): zero density
    on mx (x my) nodes."""
    return FluidSpecies(name, mu, q, m, np.zeros((mx,) if my is None else (mx, my)), T)


class DensitySource:
    """DensitySource{D} (pic/sources.jl:3-6): init adds delta (per node) to
    a fluid's density (sample!, sources.jl:36-38)."""

    def __init__(self, delta, grid):
        self.delta = np.asarray(delta, dtype=np.float64)
        self.grid = grid
        self.species = None


class MaxwellianSource:
    """MaxwellianSource{D,V} (pic/sources.jl:8-22): rate per second, position
    box wx (+ offset dx), velocity spread wv (+ drift dv)."""

    def __init__(self, rate, wx, wv, dx=None, dv=None):
        self.rate = float(rate)
        self.wx = np.atleast_1d(np.asarray(wx, dtype=np.float64)).ravel()
        self.wv = np.atleast_1d(np.asarray(wv, dtype=np.float64)).ravel()
        self.dx = np.zeros_like(self.wx) if dx is None else np.atleast_1d(
            np.asarray(dx, dtype=np.float64)).ravel()
        self.dv = np.zeros(3) if dv is None else np.atleast_1d(
            np.asarray(dv, dtype=np.float64)).ravel()
        self.species = None
        self.seed_stream = None


def thermal_speed(T, m):
    """thermal_speed (problem/configuration.jl:79-81): sqrt(2 kB T / m)."""
    return math.sqrt(2.0 * KB * T / m)


def create_kinetic_species(name, N, q, m, weight, D=2, V=3):
    """create_kinetic_species (problem/configuration.jl:95-102)."""
    s = KineticSpecies(name, N, D, V)
    s.q = float(q)
    s.m = float(m)
    s.wg *= weight
    s.w0 = float(weight)
    return s


def create_thermalized_beam(species, x, vb, dx=None, T=300.0, rate=1.0):
    """create_thermalized_beam (problem/configuration.jl:89-93)."""
    vth = thermal_speed(T, species.m) * np.ones(3)
    x = np.atleast_1d(np.asarray(x, dtype=np.float64)).ravel()
    src = MaxwellianSource(rate, x, vth, dx=np.zeros_like(x) if dx is None else dx,
                           dv=np.asarray(vb, dtype=np.float64).ravel())
    src.species = species
    return src


class BorisPusher:
    """BorisPusher{:xy} / BorisPusher{:rz} (pic/pushers.jl:4-17); E only
    (B = 0 on this path).  The :rz pusher follows the Cartesian step by
    transform_from_cartesian_to_cylindrical! (pushers.jl:52-66)."""

    def __init__(self, kind="xy"):
        self.kind = kind


def create_boris_pusher():
    return BorisPusher("xy")


def create_axial_boris_pusher():
    return BorisPusher("rz")


# ---------------------------------------------------------------------------
# surfaces and electrodes (pic/surfaces/track.jl, pic/circuit_coupling.jl,
# problem/configuration.jl:17-76)

class Surface:
    code = 0


class PeriodicSurface(Surface):
    """A surface whose hit does nothing (hit!(::Surface), hit.jl:23-28)."""
    code = L.SURF_PERIODIC


class AbsorbingSurface(Surface):
    code = L.SURF_ABSORBING


class ReflectiveSurface(Surface):
    code = L.SURF_REFLECTIVE


def create_periodic_surface():
    return PeriodicSurface()


def create_absorbing_surface():
    return AbsorbingSurface()


def create_reflective_surface():
    return ReflectiveSurface()


class SurfaceTracker:
    """SurfaceTracker{2} (track.jl:14-19): the default surface of the grid's
    boundary edges and, in call order, the surfaces and electrodes added to
    it; the device builds the edge table from them when the configuration is
    bound."""

    def __init__(self, grid, ds):
        self.grid = grid
        self.default = ds
        self.ops = []


def create_surface_tracker(grid, ds=None):
    """create_surface_tracker(grid, ds) (track.jl:101-114)."""
    return SurfaceTracker(grid, AbsorbingSurface() if ds is None else ds)


def track_surface(st, nodes, ss):
    """track_surface!(tracker, nodes, surface) (track.jl:116-118)."""
    st.ops.append(("surface", np.asarray(nodes, dtype=bool).copy(), ss))


class FixedPotentialElectrode(Surface):
    """FixedPotentialElectrode (circuit_coupling.jl:4-8)."""

    def __init__(self, nodes, phi):
        self.nodes, self.phi, self.fixed = nodes, float(phi), True
        self._ctx, self._id = None, -1


class FloatingPotentialElectrode(Surface):
    """FloatingPotentialElectrode (circuit_coupling.jl:10-15): its surface
    charge sigma (C/m^2, initial value) collects the charge of the particles
    it absorbs."""

    def __init__(self, nodes, sigma):
        self.nodes, self.sigma0, self.fixed = nodes, float(sigma), False
        self._ctx, self._id = None, -1

    def state(self):
        """sigma, dq, phi (at the reference node) and area on the device."""
        return self._ctx.electrode_state(self._id)


def create_electrode(nodes, config, fixed=False, sigma=0.0, phi=0.0):
    """create_electrode(nodes, config; fixed, sigma, phi)
    (problem/configuration.jl:17-35): a default tracker when the
    configuration has none, then the electrode (Dirichlet nodes and an
    absorbing surface when fixed; a surface charge with the reference's
    apply_neumann rows when floating) added to it."""
    if config.grid is None or config.solver is None:
        raise ValueError("create_electrode: define the grid and the field solver first")
    if config.tracker is None:
        config.tracker = create_surface_tracker(config.grid)
    mask = np.asarray(nodes, dtype=bool).copy()
    e = FixedPotentialElectrode(mask, phi) if fixed else FloatingPotentialElectrode(mask, sigma)
    config.tracker.ops.append(("electrode", mask, e))
    return e


def _bind_tracker(ctx, st):
    ctx.tracker(st.default.code)
    for op, mask, obj in st.ops:
        m = np.ascontiguousarray(mask.T).ravel()
        if op == "surface":
            ctx.tracker_surface(m, obj.code)
        else:
            obj._id = ctx.electrode(m, fixed=obj.fixed, sigma=getattr(obj, "sigma0", 0.0),
                                    phi=getattr(obj, "phi", 0.0))
            obj._ctx = ctx


# ---------------------------------------------------------------------------
# configuration (problem/configuration.jl:5-15)

class Config:
    def __init__(self, grid=None, solver=None, pusher=None, species=None, interactions=None,
                 sources=None, tracker=None, circuit=None, cells=None):
        self.grid = grid
        self.solver = solver
        self.pusher = pusher
        self.species = list(species or [])
        self.interactions = list(interactions or [])
        self.sources = list(sources or [])
        self.tracker = tracker
        self.circuit = circuit
        self.cells = cells
        self.seed = 0
        self.device = 0
        self.ctx = None


# ---------------------------------------------------------------------------
# hooks (ParticleInCell.jl:38-41) -- replace these module attributes

def enter_loop():
    pass


def after_loop(it, t, dt):
    pass


def exit_loop():
    pass


def _default_after_push(part, grid):
    wrap(part, grid)


after_push = _default_after_push

_recording = {}


def wrap(part, grid, dims=None):
    """wrap! (pic/surfaces/wrap.jl:20-33): periodic re-entry along dims
    (1-based like the reference; default all).  Inside `after_push` it
    selects the boundary the push kernel applies."""
    _boundary(part, grid, dims, L.BND_WRAP)


def discard(part, grid, dims=None):
    """discard! (pic/surfaces/wrap.jl:1-18): particles leaving along dims are
    removed (compacted at the next sort instead of swap-removed)."""
    _boundary(part, grid, dims, L.BND_DISCARD)


def _boundary(part, grid, dims, mode):
    modes = _recording.get(id(part))
    if modes is None:
        raise RuntimeError("wrap!/discard! run inside the fused push kernel; call them from "
                           "the after_push hook, which selects the boundary per dimension")
    for d in (range(1, grid.dim + 1) if dims is None else np.atleast_1d(dims)):
        modes[int(d) - 1] = mode


# ---------------------------------------------------------------------------
# binding a configuration to a device context

def _fluid_species(config):
    return [s for s in config.species if isinstance(s, FluidSpecies)]


def _kinetic_species(config):
    return [s for s in config.species if isinstance(s, KineticSpecies)]


def _fluid_can_change(f, g):
    """The reference's update can change this fluid's density: it diffuses
    (mu != 0) or, in 2-D, is advected by E (q != 0)."""
    return f.mu != 0.0 or (g.dim == 2 and f.q != 0.0)


def _fluid_is_static(f, g):
    """True when the reference's fluid update leaves n unchanged at every step
    (advance!(::FluidSpecies), ParticleInCell.jl:74-82, with
    calculate_advection_diffusion, advection_diffusion.jl:4-38).  1-D: only
    diffusion mu (n[i-1] - 2 n[i] + n[i+1]) at interior nodes (advection is
    commented out there).  2-D: diffusion by the 5-point sum plus advection by
    v = E dt q/m, over interior nodes only, so a grid with fewer than three
    nodes along an axis has no update at all."""
    n = np.asarray(np.broadcast_to(f.n, g.n), dtype=np.float64)
    if g.dim == 1:
        if g.n[0] < 3 or f.mu == 0.0:
            return True
        return not np.any(n[:-2] - 2.0 * n[1:-1] + n[2:])
    if g.n[0] < 3 or g.n[1] < 3:
        return True
    if f.q != 0.0:
        return False
    if f.mu == 0.0:
        return True
    lap = n[:-2, 1:-1] + n[1:-1, :-2] - 4.0 * n[1:-1, 1:-1] + n[2:, 1:-1] + n[1:-1, 2:]
    return not np.any(lap)


def bind(config):
    """Create the device context for `config` and move grid, solver,
    species, fluids and interactions onto it.  Called by `solve`; callable
    directly to use the operators below outside the loop.  Fluids the
    reference's update cannot change join the constant background; the
    others are advanced on the device every step.  An MCC target gas whose
    density the update would change is refused with ISKRA_E_UNSUPPORTED
    (the collisions read a fixed target density)."""
    from ._lib import Context, IskraError
    from . import chemistry

    g = config.grid
    if not isinstance(g, UniformGrid):
        raise TypeError("config.grid must come from create_uniform_grid")
    targets = {id(i.target) for i in config.interactions
               if isinstance(i, chemistry.MonteCarloCollisions)}
    mobile = []
    for f in _fluid_species(config):
        f._ctx = None
        if id(f) in targets:
            if not _fluid_is_static(f, g):
                raise IskraError(L.E_UNSUPPORTED,
                                 f"{f.name}: an MCC target gas whose density would change "
                                 "every step (advection_diffusion.jl:4-38); the collisions "
                                 "read a fixed target density")
        elif _fluid_can_change(f, g):
            mobile.append(f)
    ctx = Context(config.device)
    _grid_to(ctx, g)
    if config.pusher is not None and getattr(config.pusher, "kind", "xy") == "rz":
        ctx.pusher(L.PUSHER_RZ)
    elif config.pusher is not None:
        ctx.pusher(L.PUSHER_XY)
    ctx.rng(config.seed)
    ctx._seed = config.seed
    config.ctx = ctx
    for s in _kinetic_species(config):
        if s.D != g.dim:
            raise ValueError(f"{s.name}: D = {s.D} on a {g.dim}-D grid")
        s._ctx = None
        s._pull()
        s._sid = ctx.species(s.q, s.m, s.w0, s.N, track_ids=True)
        modes = [L.BND_NONE, L.BND_NONE]
        _recording[id(s)] = modes
        try:
            after_push(s, g)
        finally:
            del _recording[id(s)]
        ctx.boundary(s._sid, modes[:g.dim])
        s._ctx = ctx
        s._dev_stale = True
        s._push()
    bg = np.zeros(g.n)
    for f in _fluid_species(config):
        if f not in mobile:
            bg += f.q * np.broadcast_to(f.n, g.n)
    ctx.set_field(L.F_BG, bg.T.ravel())
    for f in mobile:
        f._n = np.array(np.broadcast_to(f._n, g.n), dtype=np.float64)
        f._grid_n = tuple(g.n)
        f._fid = ctx.fluid(f.q, f.m, f.mu, np.ascontiguousarray(f._n.T).ravel())
        f._ctx = ctx
        f._host_stale = f._dev_stale = False
    if isinstance(config.solver, PoissonSolver):
        config.solver._ctx = ctx
        config.solver._configure(ctx)
    if config.tracker is not None:
        _bind_tracker(ctx, config.tracker)
    if config.circuit is not None:
        config.circuit._bind(ctx)
    for inter in config.interactions:
        if isinstance(inter, (chemistry.MonteCarloCollisions,
                              chemistry.DirectSimulationMonteCarlo)):
            inter._bind(ctx, g)
    diagnostics.register_simulation(config)
    return ctx


def _grid_to(ctx, g):
    """The grid on the device: create_uniform_grid or create_axial_grid."""
    if g.kind == "rz":
        vp = g.vol_periodic()
        ctx.grid_axial(g.n, g.dh, g.origin, (vp[2], vp[3]))
    else:
        ctx.grid(g.dim, g.n, g.dh, g.origin if g.dim == 2 else (g.origin[0], 0.0),
                 g.vol_periodic())


def _sync_to_device(config):
    for s in _kinetic_species(config):
        s._push()
    for f in _fluid_species(config):
        f._push()


def _mark_device_newer(config):
    for s in _kinetic_species(config):
        s._host_stale = True
        s._dev_stale = False
    for f in _fluid_species(config):
        if f._ctx is not None:
            f._host_stale = True
            f._dev_stale = False


# ---------------------------------------------------------------------------
# operators (each one a device call)

def init(src, species, dt):
    """init(src, species, dt) = sample!(src, species, dt) (ParticleInCell.jl:47,
    pic/sources.jl:24-34): append min(free capacity, floor(rate*dt))
    particles drawn on the device (Philox, keyed by the species' source
    count).  The draw stream is (species << 24) | call count: species below
    64 and up to 2^24 calls per species keep every call's stream distinct,
    and below the MCC streams (0x4D43xxxx)."""
    if isinstance(src, DensitySource):
        # sample!(::DensitySource, ::FluidSpecies) (sources.jl:36-38): n .+= delta
        if not isinstance(species, FluidSpecies):
            raise TypeError("init: a DensitySource feeds a FluidSpecies")
        n = species.n
        n += np.broadcast_to(src.delta, n.shape)
        return
    if species._ctx is None:
        raise RuntimeError("init: the species is not bound to a simulation yet (call it from "
                           "enter_loop, or bind(config) first)")
    species._push()
    n = int(math.floor(src.rate * dt))
    ctx = species._ctx
    count = getattr(species, "_nsample", 0)
    if species._sid >= 64 or count >= 1 << 24:
        raise L.IskraError(L.E_UNSUPPORTED,
                           f"{species.name}: sample! call {count} of species {species._sid} has "
                           "no distinct draw stream (at most 2^24 calls for 64 species)")
    species._nsample = count + 1
    wx = src.wx[:species.D]
    dx = src.dx[:species.D]
    ctx.sample_maxwellian(species._sid, n, wx, src.wv, dx=dx, dv=src.dv,
                          seed=ctx_seed(ctx), stream=(species._sid << 24) | count)
    species._host_stale = True


def ctx_seed(ctx):
    return getattr(ctx, "_seed", 0)


def advance(part, E, B, dt, config):
    """advance!(::KineticSpecies, E, B, dt, config) (ParticleInCell.jl:51-61):
    gather E, Boris push (B = 0), after_push boundary, and this species'
    CIC deposit, in one kernel, with the device's current E."""
    part._push()
    config.ctx.push(part._sid, dt)
    part._host_stale = True


def push_particles(pusher, part, E, B, dt):
    """push_particles! (pic/pushers.jl:8-11) is fused with the gather, the
    boundary and the deposit; use advance."""
    raise RuntimeError("push_particles!: the push is fused with gather, boundary and deposit; "
                       "call advance(part, E, B, dt, config)")


def particle_to_grid(part, grid):
    """particle_to_grid(part, grid, p -> wg[p]) (pic/cloud_in_cell.jl:1-18):
    CIC sum of the particle weights per node."""
    from .regular_grids import cell_volume
    return density(part, grid) * cell_volume(grid)


def density(part, grid):
    """density(::KineticSpecies, grid) (pic/kinetic.jl:53): CIC weights per
    node divided by cell_volume; FluidSpecies return n (fluid.jl:11)."""
    if isinstance(part, FluidSpecies):
        return part.n
    part._push()
    ctx = part._ctx
    ctx.deposit_species(part._sid)
    ctx.deposit()
    return ctx.field(L.F_DENSITY, part._sid).reshape(grid.n[::-1]).T.copy()


def grid_to_particle(grid, part, E=None):
    """grid_to_particle(grid, part, (i,j) -> E[i,j,:]) (pic/cloud_in_cell.jl:20-36):
    bilinear interpolation of E to the particles (np x 3).  E (nx[,ny],3)
    replaces the device field first when given."""
    part._push()
    ctx = part._ctx
    if E is not None:
        E = np.asarray(E, dtype=np.float64)
        ctx.set_field(L.F_EX, E[..., 0].T.ravel())
        if grid.dim == 2:
            ctx.set_field(L.F_EY, E[..., 1].T.ravel())
    return ctx.gather(part._sid)


# ---------------------------------------------------------------------------
# the loop (ParticleInCell.jl:84-139)

_noop_after_loop = after_loop


def checkpoint(config, path):
    """Save the bound simulation's state to path (.npz, no pickled objects):
    the device state (iskra_checkpoint_save), the sources' call counters and
    the step count.  No reference counterpart (its TODO:3)."""
    _sync_to_device(config)
    ctx = config.ctx
    np.savez(path, blob=ctx.checkpoint(),
             nsample=np.array([getattr(s, "_nsample", 0) for s in _kinetic_species(config)],
                              dtype=np.int64),
             step=np.array([ctx.stats()["steps"] + getattr(ctx, "_step0", 0)], dtype=np.int64))


def _restore(config, path):
    with np.load(path, allow_pickle=False) as z:
        blob, ns, step = z["blob"], z["nsample"], int(z["step"][0])
    config.ctx.restore(blob)
    config.ctx._step0 = step
    for s, n in zip(_kinetic_species(config), ns):
        s._nsample = int(n)
    _mark_device_newer(config)
    return step


def solve(config, dt=1e-5, timesteps=200, seed=None, resume=None):
    """solve(config, dt, timesteps) (ParticleInCell.jl:84-139).

    Per iteration: sources, MCC and DSMC interactions (with the previous E),
    advance! of every kinetic species (gather, push, after_push boundary,
    deposit) and of every mobile fluid, rho = sum q n, phi from the Poisson
    solve, E, then after_loop(it, it*dt - dt, dt).  When after_loop is the
    default and there are no per-step sources, the iterations run as one
    device call (two with fluid species, whose drift-velocity record needs
    the E the last step starts from, see iskra_amd.diagnostics).

    resume: a file written by checkpoint(); the configuration must be set
    up as the saved one was.  Its state replaces enter_loop, and the
    iterations continue the saved run's numbering; the results equal those
    of the run that was never interrupted."""
    if seed is not None:
        config.seed = int(seed)
    ctx = bind(config)
    it0 = 0
    if resume is not None:
        it0 = _restore(config, resume)
    else:
        enter_loop()
    _sync_to_device(config)
    dt = float(dt)
    fused = after_loop is _noop_after_loop and not config.sources
    if fused:
        n = int(timesteps)
        if n > 1 and diagnostics.needs_step_begin(config):
            ctx.step(dt, n - 1)
            n = 1
        diagnostics.step_begin(config, dt)
        if n > 0:
            ctx.step(dt, n)
        _mark_device_newer(config)
    else:
        for it in range(it0 + 1, it0 + int(timesteps) + 1):
            for src in config.sources:
                init(src, src.species, dt)
            _sync_to_device(config)
            diagnostics.step_begin(config, dt)
            ctx.step(dt, 1)
            _mark_device_newer(config)
            after_loop(it, it * dt - dt, dt)
            _sync_to_device(config)
    exit_loop()
    return ctx


def field(config, name):
    """Node field of the bound simulation: "rho", "phi", "E" (nx[,ny],3)."""
    ctx = config.ctx
    n = config.grid.n
    if name == "rho":
        return ctx.field(L.F_RHO).reshape(n[::-1]).T.copy()
    if name == "phi":
        return ctx.field(L.F_PHI).reshape(n[::-1]).T.copy()
    if name == "E":
        E = np.zeros(n + (3,))
        E[..., 0] = ctx.field(L.F_EX).reshape(n[::-1]).T
        if len(n) == 2:
            E[..., 1] = ctx.field(L.F_EY).reshape(n[::-1]).T
        return E
    raise KeyError(name)
