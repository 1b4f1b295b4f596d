"""Diagnostics (Diagnostics/src/Diagnostics.jl, hdf5.jl, openpmd/*.jl): the
registry of field, particle and probe records and their openPMD files.

    import iskra_amd.diagnostics as diag
    def iteration(i, t, dt):
        if i % 100 == 0:
            diag.new_iteration(lambda it: (diag.save_records(it, "e-/"),
                                           diag.save_record(it, "rho")),
                               "run", i, t, dt)
    pic.after_loop = iteration

writes run/hdf5/data<i>.h5 with /data/<i>/fields/rho and
/data/<i>/particles/e-/{id,mass,charge,weighting,momentum,position,
positionOffset} as the reference's new_iteration do-block does
(hdf5.jl:47-101).  The files are written by libiskra_pmd.so (libhdf5).

The reference copies every registered array into its record at every step
(@particle after each push, ParticleInCell.jl:63-71; @field after the
deposit and the field solve, :123,130-133; the collision frequencies,
mcc.jl:287 and dsmc.jl:141; the circuit probes, Circuit.jl:132-135).  Here
binding a simulation registers the same keys as records that read the device
when a file is written, so steps that write nothing copy nothing.  What a
file holds is what the reference's records hold after the step:
  - particles, densities, rho, phi, E, B and the probes are the state at the
    end of the step (nothing later in the reference's step changes them);
  - nuMCC-<species>-<k> and nuDSMC count the events of the last step (the
    device sets the earlier steps' counts apart before the last step of
    every step call);
  - the fluid records hold n, the last step's dn (zero for a fluid the
    update cannot change) and v = E dt q/m with the E of the step's start,
    as the reference's advance! computes them (ParticleInCell.jl:74-82).
Particles are written in the device's order; the id record pairs them with
the reference's.  Records a script registers itself with `field`,
`particle` or `probe` hold a copy of the array passed, as the reference's
update! does.
"""
import ctypes as C
import math
import os

import numpy as np

from . import _lib as L

# ---------------------------------------------------------------------------
# the writer library (include/iskra_pmd.h)

PMD_LIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libiskra_pmd.so")
PMD_F64, PMD_U32, PMD_I64 = 0, 1, 2
PMD_EXPORTS = ["iskra_pmd_open", "iskra_pmd_close", "iskra_pmd_units", "iskra_pmd_field",
               "iskra_pmd_particle", "iskra_pmd_particle_constant", "iskra_pmd_probe",
               "iskra_pmd_last_error"]


class Root(C.Structure):
    _fields_ = [("author", C.c_char_p), ("seed", C.c_int64), ("software", C.c_char_p),
                ("software_version", C.c_char_p), ("date", C.c_char_p)]


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int32), ("axial", C.c_int32), ("nodes", C.c_int64 * 2),
                ("spacing", C.c_double * 2)]


_PMD = None


def pmd_lib():
    """libiskra_pmd.so, loaded on first use; raises if it was not built."""
    global _PMD
    if _PMD is None:
        if not os.path.exists(PMD_LIBPATH):
            raise RuntimeError(f"{PMD_LIBPATH} is missing: build it with python -m iskra_amd.build "
                               "(it needs the HDF5 headers and library)")
        lib = C.CDLL(PMD_LIBPATH)
        P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        S = C.c_char_p
        sig = {
            "iskra_pmd_open": (I, [S, I64, D, D, C.POINTER(Root), C.POINTER(P)]),
            "iskra_pmd_close": (I, [P]),
            "iskra_pmd_units": (I, [S, P, P]),
            "iskra_pmd_field": (I, [P, S, S, C.POINTER(Grid), I, I, P, P, D]),
            "iskra_pmd_particle": (I, [P, S, S, I64, I, I, I, P, I64, I, D]),
            "iskra_pmd_particle_constant": (I, [P, S, S, D, I64, I, D]),
            "iskra_pmd_probe": (I, [P, S, S, D, D]),
            "iskra_pmd_last_error": (S, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _PMD = lib
    return _PMD


def _check(rc):
    if rc != 0:
        raise L.IskraError(rc, pmd_lib().iskra_pmd_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def units(u):
    """usi(units) (openpmd/root.jl:17-38): (unitDimension as 7 floats, unitSI)."""
    dims = np.zeros(7)
    si = np.zeros(1)
    _check(pmd_lib().iskra_pmd_units(u.encode(), _ptr(dims), _ptr(si)))
    return tuple(dims), float(si[0])


# ---------------------------------------------------------------------------
# records (openpmd/fields.jl, particles.jl, circuit.jl)

def _grid_struct(grid):
    g = Grid()
    g.dim = grid.dim
    g.axial = 1 if getattr(grid, "kind", "xy") == "rz" else 0
    for k in range(grid.dim):
        g.nodes[k] = int(grid.n[k])
        g.spacing[k] = float(grid.dh[k])
    return g


def _planes(data, grid, withcomponents):
    """A reference-shaped node array (nx[,ny][,K]) as planes with x fastest:
    (K, nodes) for withcomponents, else (1, nodes) of the first plane
    (view(input, :, :, 1), fields.jl:47-49)."""
    a = np.asarray(data, dtype=np.float64)
    n = tuple(int(v) for v in grid.n)
    if a.shape[:len(n)] != n:
        raise ValueError(f"field data of shape {a.shape} on a grid of {n} nodes")
    a = a.reshape(n + (-1,))
    K = a.shape[-1] if withcomponents else 1
    if K > 3:
        raise ValueError("a field record has at most 3 components")
    return np.stack([a[..., k].ravel(order="F") for k in range(K)])


class FieldRecord:
    """FieldRecord (openpmd/fields.jl:15-63): `source()` gives (K, nodes)
    planes with x fastest."""

    def __init__(self, units, grid, source, withcomponents=False, offset=0.0, pos=None):
        self.units, self.grid, self.source = units, grid, source
        self.withcomponents = bool(withcomponents)
        self.offset = float(offset)
        self.pos = None if pos is None else np.asarray(pos, dtype=np.float64)

    def save(self, it, key):
        planes = np.ascontiguousarray(self.source(), dtype=np.float64)
        g = _grid_struct(self.grid)
        pos = None if self.pos is None else _ptr(np.ascontiguousarray(self.pos))
        _check(pmd_lib().iskra_pmd_field(it._h, key.encode(), self.units.encode(), C.byref(g),
                                         planes.shape[0], int(self.withcomponents),
                                         _ptr(planes), pos, self.offset))


class ParticleRecord:
    """ParticleRecord (openpmd/particles.jl:10-51): `source(cache)` gives
    (np, value) for a single-value record, else (np, (ncomp, np) array)."""

    def __init__(self, units, source, weighted=False, withcomponents=False, offset=0.0):
        self.units, self.source = units, source
        self.weighted = bool(weighted)
        self.withcomponents = bool(withcomponents)
        self.offset = float(offset)

    def save(self, it, key):
        np_, val = self.source(it._cache)
        lib = pmd_lib()
        if np.ndim(val) == 0:
            _check(lib.iskra_pmd_particle_constant(it._h, key.encode(), self.units.encode(),
                                                   float(val), int(np_), int(self.weighted),
                                                   self.offset))
            return
        a = np.asarray(val)
        dtype = {np.dtype(np.uint32): PMD_U32, np.dtype(np.int64): PMD_I64}.get(a.dtype, PMD_F64)
        a = np.ascontiguousarray(a, dtype=a.dtype if dtype != PMD_F64 else np.float64)
        if a.ndim == 1:
            a = a[None, :]
        ncomp = a.shape[0] if self.withcomponents else 1
        _check(lib.iskra_pmd_particle(it._h, key.encode(), self.units.encode(), int(np_), ncomp,
                                      int(self.withcomponents), dtype, _ptr(a), int(np_),
                                      int(self.weighted), self.offset))


class CircuitProbeRecord:
    """CircuitProbeRecord (Diagnostics/src/circuit.jl:1-19): `source()`
    gives the probed value."""

    def __init__(self, units, source, offset=0.0):
        self.units, self.source, self.offset = units, source, float(offset)

    def save(self, it, key):
        _check(pmd_lib().iskra_pmd_probe(it._h, key.encode(), self.units.encode(),
                                         float(self.source()), self.offset))


records = {}  # the registry (Diagnostics.jl:22)


def register_diagnostic(key, record):
    """register_diagnostic (Diagnostics.jl:24-31) for an already built record:
    the first registration of a key prints as the reference does; a later one
    replaces the record (the reference's update!)."""
    if key not in records:
        print("Registering diagnostics", key)
    records[key] = record


def field(key, units, data, grid, offset=0.0, pos=None, withcomponents=False):
    """@field key units data grid (Diagnostics.jl:33-38): a copy of data
    (nx[,ny][,K] like the reference's arrays)."""
    planes = _planes(data, grid, withcomponents)
    register_diagnostic(key, FieldRecord(units, grid, lambda: planes, withcomponents, offset, pos))


def particle(key, units, data, part, offset=0.0, weighted=False, withcomponents=False):
    """@particle key units data part (Diagnostics.jl:40-45): a copy of the
    first part.np rows of data (N or N x M); a one-element data is a
    single-value record of shape (np,)."""
    a = np.array(data)
    n = int(part.np)
    if a.size == 1:
        v = a.ravel()[0]
        rec = ParticleRecord(units, lambda cache: (n, v), weighted, withcomponents, offset)
    else:
        a = a.reshape(a.shape[0], -1)[:n]
        cols = np.ascontiguousarray(a.T if withcomponents else a[:, :1].T)
        rec = ParticleRecord(units, lambda cache: (n, cols), weighted, withcomponents, offset)
    register_diagnostic(key, rec)


def probe(key, data, units, offset=0.0):
    """@probe key data units (Diagnostics/src/circuit.jl:22-27)."""
    v = float(np.asarray(data, dtype=np.float64).ravel()[0])
    register_diagnostic(key, CircuitProbeRecord(units, lambda: v, offset))


# ---------------------------------------------------------------------------
# files (hdf5.jl:47-101)

class Iteration:
    """The open /data/<i> group of new_iteration's do-block."""

    def __init__(self, prefix, i, t, dt, seed=0, author=None):
        self._h = None
        self._cache = {}
        root = Root(author=None if author is None else author.encode(), seed=int(seed))
        h = C.c_void_p()
        _check(pmd_lib().iskra_pmd_open(os.fspath(prefix).encode(), int(i), float(t), float(dt),
                                        C.byref(root), C.byref(h)))
        self._h = h
        self.path = os.path.join(os.fspath(prefix), "hdf5", f"data{int(i)}.h5")

    def close(self):
        if self._h is not None:
            h, self._h = self._h, None
            self._cache.clear()
            _check(pmd_lib().iskra_pmd_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def new_iteration(f, prefix, i, t, dt, seed=0, author=None):
    """new_iteration(f, prefix, i, t, dt) (hdf5.jl:47-66): open
    <prefix>/hdf5/data<i>.h5, call f(it) and close the file even if f
    raises.  Returns the file's path."""
    it = Iteration(prefix, i, t, dt, seed=seed, author=author)
    try:
        f(it)
    finally:
        it.close()
    return it.path


def save_record(it, key, record=None):
    """save_record(it, key) (Diagnostics.jl:56-62, hdf5.jl:69-101); an
    unknown key prints the reference's message and writes nothing."""
    if record is None:
        record = records.get(key)
        if record is None:
            print("Couldn't find diagnostic", key)
            return
    if it._h is None:
        raise L.IskraError(L.E_STATE, f"save_record {key}: the iteration file is closed")
    record.save(it, key)


def save_records(it, prefix):
    """save_records(it, prefix) (Diagnostics.jl:48-54): every record whose
    key starts with prefix."""
    for key in list(records):
        if key.startswith(prefix):
            save_record(it, key, records[key])


# ---------------------------------------------------------------------------
# the records a bound simulation registers (the keys of the reference's loop)

def _species_state(part, cache):
    """(np, x (np x D), v (np x 3), id) of a kinetic species, downloaded once
    per file.  Reads through the species' host mirror without marking it
    changed, so writing a file never re-uploads or reorders particles."""
    got = cache.get(id(part))
    if got is None:
        part._pull()
        n = int(part._np)
        got = (n, part._x[:n].copy(), part._v[:n].copy(), part._id[:n].copy())
        cache[id(part)] = got
    return got


def _register_species(part):
    def xs(cache):
        n, x, _, _ = _species_state(part, cache)
        return n, x.T
    def mom(cache):
        n, _, v, _ = _species_state(part, cache)
        return n, (part.m * v).T
    def ids(cache):
        n, _, _, i = _species_state(part, cache)
        return n, i[None, :]
    def wg(cache):
        n = _species_state(part, cache)[0]
        return n, np.full((1, n), part.w0)
    def const(value):
        return lambda cache: (_species_state(part, cache)[0], value)
    nm = part.name
    records[nm + "/id"] = ParticleRecord("1", ids)
    records[nm + "/mass"] = ParticleRecord("kg", const(part.m))
    records[nm + "/charge"] = ParticleRecord("C", const(part.q))
    records[nm + "/weighting"] = ParticleRecord("1", wg, weighted=True)
    records[nm + "/momentum"] = ParticleRecord("kg*m/s", mom, withcomponents=True)
    records[nm + "/position"] = ParticleRecord("m", xs, withcomponents=True)
    for ax in "xyz":
        records[f"{nm}/positionOffset/{ax}"] = ParticleRecord("m", const(0.0))


def _efield(ctx, grid):
    E = np.zeros((3, ctx.nn))
    E[0] = ctx.field(L.F_EX)
    if grid.dim == 2:
        E[1] = ctx.field(L.F_EY)
    return E


class _LoopState:
    """What the loop keeps for the records: the step's dt and the E of the
    step's start (the fluids' drift velocity)."""

    def __init__(self):
        self.dt = 0.0
        self.E_start = None


def register_simulation(config):
    """Register the records of the reference's loop for a bound
    configuration (called by particle_in_cell.bind)."""
    from . import chemistry
    from .particle_in_cell import FluidSpecies, KineticSpecies

    ctx, g = config.ctx, config.grid
    state = _LoopState()
    config._diag = state
    kinetic = [s for s in config.species if isinstance(s, KineticSpecies)]
    for part in kinetic:
        _register_species(part)
        records["n" + part.name] = FieldRecord(
            "1/m^2", g, (lambda p: lambda: ctx.field(L.F_DENSITY, p._sid)[None, :])(part))
    for f in (s for s in config.species if isinstance(s, FluidSpecies)):
        # density(fluid) = n enters the density loop's records too (ParticleInCell.jl:118-124)
        if f._ctx is not None:  # advanced on the device
            records["n" + f.name] = FieldRecord(
                "1/m^2", g, (lambda f: lambda: ctx.field(L.F_FLUID_N, f._fid)[None, :])(f))
            records["d" + f.name] = FieldRecord(
                "1/m^2", g, (lambda f: lambda: ctx.field(L.F_FLUID_DN, f._fid)[None, :])(f))
        else:
            nconst = np.ascontiguousarray(
                np.broadcast_to(f._n, g.n), dtype=np.float64).T.ravel()[None, :]
            records["n" + f.name] = FieldRecord("1/m^2", g, lambda c=nconst: c)
            zero = np.zeros((1, ctx.nn))
            records["d" + f.name] = FieldRecord("1/m^2", g, lambda z=zero: z)
        def vel(f=f):
            E = state.E_start if state.E_start is not None else np.zeros((3, ctx.nn))
            return E * (state.dt * (f.q / f.m))
        records["v" + f.name] = FieldRecord("m/s", g, vel, withcomponents=True)
    records["rho"] = FieldRecord("C/m^2", g, lambda: ctx.field(L.F_RHO)[None, :])
    records["phi"] = FieldRecord("V", g, lambda: ctx.field(L.F_PHI)[None, :])
    records["E"] = FieldRecord("V/m", g, lambda: _efield(ctx, g), withcomponents=True)
    records["B"] = FieldRecord("A/m", g, lambda: np.zeros((3, ctx.nn)), withcomponents=True)
    for inter in config.interactions:
        if isinstance(inter, chemistry.MonteCarloCollisions):
            for k in range(len(inter.collisions)):
                records[f"nuMCC-{inter.source.name}-{k + 1}"] = FieldRecord(
                    "1/m^2", g, (lambda m, k: lambda: m._last_step()[k][None, :].astype(
                        np.float64))(inter, k))
        elif isinstance(inter, chemistry.DirectSimulationMonteCarlo):
            records["nuDSMC"] = FieldRecord(
                "1/m^2", g, (lambda d: lambda: d._last_step()[None, :].astype(np.float64))(inter))
    cir = config.circuit
    if cir is not None:
        def v1():
            t = cir.state()["t"] - state.dt  # the source is probed at the step's start time
            V = cir.V
            return V.offset + V.amplitude * math.sin(2.0 * math.pi * V.freq * t + V.phase)
        records["Q1"] = CircuitProbeRecord("C", lambda: cir.state()["q"])
        records["I1"] = CircuitProbeRecord("A", lambda: cir.state()["i"])
        records["V1"] = CircuitProbeRecord("V", v1)
        records["Vext"] = CircuitProbeRecord("V", lambda: cir.state()["vext"])


def needs_step_begin(config):
    """True when some record depends on where the last step starts (the
    fluids' drift velocity uses the E of that moment)."""
    from .particle_in_cell import FluidSpecies

    return any(isinstance(s, FluidSpecies) for s in config.species)


def step_begin(config, dt):
    """Called by the loop before each device step that may be saved: keep
    the step's dt and, with fluid species, the E the step starts from."""
    from .particle_in_cell import FluidSpecies

    state = getattr(config, "_diag", None)
    if state is None:
        return
    state.dt = float(dt)
    if any(isinstance(s, FluidSpecies) for s in config.species):
        state.E_start = _efield(config.ctx, config.grid)
