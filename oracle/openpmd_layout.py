"""Expected openPMD layout of the reference's diagnostics files (test
infrastructure: the tests compare files written by libiskra_pmd.so, read
back with h5py, against this).

Restates what Diagnostics/src/hdf5.jl:47-101 writes with HDF5.jl 0.15.4
(Manifest.toml:118-122) for the records of Diagnostics/src/openpmd/
root.jl, fields.jl, particles.jl and circuit.jl.  HDF5.jl's conventions
for Julia values, which fix the on-disk types:
  - String: fixed-length UTF-8 string of its byte length, scalar
    dataspace; "" has a null dataspace;
  - NTuple of Strings (setindex!, hdf5.jl:23-24, writes [val...]): 1-D
    array of variable-length UTF-8 strings;
  - NTuple of Chars (hdf5.jl:5-19): 1-D array of one-byte UTF-8 strings;
  - Float64 / Int64 / UInt32 / UInt8 scalars and NTuple{N,Float64} or a
    Dims tuple (as [val...]) as native scalars and 1-D arrays;
  - an array of size (n1, n2) is a dataset of dims (n2, n1) (Julia is
    column-major), so a node field of an (nx, ny) grid has dims (ny, nx) with
    x varying fastest and a 1-D field, written as view(input, :, :, 1) of a
    vector (fields.jl:47-49), has dims (1, nx).
Attribute descriptors are those of tests/pmd_dump.py.

One departure is deliberate and stated where it applies: the reference's
CircuitProbeRecord passes ("","") for the NTuple{2,Char} axisLabels
(circuit.jl:9-15), which Julia cannot convert; the probe records here hold
two empty one-byte strings there.
"""

# unitDimension (L, M, T, I, theta, N, J) of the unit strings the
# reference's loop and collision code register (ParticleInCell.jl:63-82,
# 123, 130-133; mcc.jl:287; dsmc.jl:141; Circuit.jl:132-135), from the SI
# definitions of the units (C = A s, V = kg m^2 s^-3 A^-1)
UNIT_DIMENSION = {
    "1": (0, 0, 0, 0, 0, 0, 0),
    "kg": (0, 1, 0, 0, 0, 0, 0),
    "m": (1, 0, 0, 0, 0, 0, 0),
    "C": (0, 0, 1, 1, 0, 0, 0),
    "A": (0, 0, 0, 1, 0, 0, 0),
    "V": (2, 1, -3, -1, 0, 0, 0),
    "kg*m/s": (1, 1, -1, 0, 0, 0, 0),
    "m/s": (1, 0, -1, 0, 0, 0, 0),
    "1/m^2": (-2, 0, 0, 0, 0, 0, 0),
    "C/m^2": (-2, 0, 1, 1, 0, 0, 0),
    "V/m": (1, 1, -3, -1, 0, 0, 0),
    "A/m": (-1, 0, 0, 1, 0, 0, 0),
}


def s(v):
    """a Julia String attribute"""
    b = v.encode("utf-8")
    return ["str", v, len(b)] if b else ["null_str"]


def f64(v):
    return ["f64", float(v)]


def f64v(v):
    return ["f64v", [float(x) for x in v]]


def root_attrs(seed=0, author="", software="iskra", version="git+?"):
    """RootMetadata (root.jl:46-63); the date is checked for its format only."""
    return {
        "openPMD": s("1.1.0"), "openPMDextension": ["u32", 1],
        "iterationEncoding": s("fileBased"), "iterationFormat": s("data%T.h5"),
        "basePath": s("/data/%T"), "meshesPath": s("fields/"),
        "particlesPath": s("particles/"), "author": s(author), "seed": ["i64", int(seed)],
        "software": s(software), "softwareVersion": s(version),
    }


def meshes_attrs():
    """FieldMetadata() (fields.jl:1-12) on /data/<i>/fields"""
    return {
        "fieldSolver": s("other"), "fieldSolverParameters": s("Nagel"),
        "fieldBoundary": ["vstr", ["open"] * 4],
        "particleBoundary": ["vstr", ["periodic"] * 4],
        "currentSmoothing": s("none"), "chargeCorrection": s("none"),
    }


def iteration_attrs(t, dt):
    """Iteration(dt, t, 1.0) (hdf5.jl:48, root.jl:40-44) on /data/<i>"""
    return {"dt": f64(dt), "time": f64(t), "timeUnitSI": f64(1.0)}


def particles_attrs():
    """ParticleMetadata() (particles.jl:1-8) on /data/<i>/particles"""
    return {
        "particleShape": f64(0.0), "currentDeposition": s("none"), "particlePush": s("Boris"),
        "particleInterpolation": s("other"), "particleSmoothing": s("none"),
    }


def skeleton(i, t, dt, seed=0, author=""):
    """The objects new_iteration creates (hdf5.jl:47-59): path -> entry."""
    base = f"/data/{i}"
    return {
        "/": {"kind": "group", "attrs": root_attrs(seed, author)},
        "/data": {"kind": "group", "attrs": {}},
        base: {"kind": "group", "attrs": iteration_attrs(t, dt)},
        base + "/fields": {"kind": "group", "attrs": meshes_attrs()},
        base + "/particles": {"kind": "group", "attrs": particles_attrs()},
    }


def _ensure_groups(tree, path):
    """intermediate groups HDF5.jl creates for a path (no attributes)"""
    parts = path.strip("/").split("/")
    for k in range(1, len(parts)):
        p = "/" + "/".join(parts[:k])
        tree.setdefault(p, {"kind": "group", "attrs": {}})


def field_record(tree, i, key, units, grid, planes, withcomponents, unit_si=1.0,
                 position=None, offset=0.0, probe=False):
    """save_record(it, key, ::FieldRecord) (hdf5.jl:86-95) of a record built
    by init!(FieldRecord) (fields.jl:27-63), or of a CircuitProbeRecord
    (hdf5.jl:97-101, circuit.jl:1-19) when probe is set.

    grid: dict(dim, axial, nodes, spacing).  planes: list of node planes as
    flat lists with x fastest (one for a plain record, the x, y, z planes for
    withcomponents).  Returns {path: (dims, flat values)} of the datasets."""
    f = f"/data/{i}/fields/{key}"
    if probe:
        L, labels = 2, ["", ""]
        spacing, dims = [0.0, 0.0], (1, 1)
    else:
        L = grid["dim"]
        labels = list("rz"[:L] if grid["axial"] else "xy"[:L])
        spacing = list(grid["spacing"][:L])
        nx = grid["nodes"][0]
        ny = grid["nodes"][1] if L == 2 else 1
        dims = (ny, nx)
    pos = [0.0] * L if position is None else list(position)
    comp = {"position": f64v(pos), "unitSI": f64(unit_si)}
    rec = {
        "unitDimension": f64v(UNIT_DIMENSION[units]), "timeOffset": f64(offset),
        "axisLabels": ["chars", labels], "dataOrder": s("C"), "geometry": s("cartesian"),
        "geometryParameters": s(""), "gridGlobalOffset": f64v([0.0] * L),
        "gridSpacing": f64v(spacing), "gridUnitSI": f64(1.0), "fieldSmoothing": s("none"),
    }
    out = {}
    _ensure_groups(tree, f)
    if not withcomponents:
        tree[f] = {"kind": "dataset", "dtype": "<f8", "shape": list(dims),
                   "attrs": {**comp, **rec}}
        out[f] = (dims, planes[0])
    else:
        tree[f] = {"kind": "group", "attrs": rec}
        n = dims[0] * dims[1]
        for k, ax in enumerate("xyz"):
            g = f + "/" + ax
            tree[g] = {"kind": "dataset", "dtype": "<f8", "shape": list(dims),
                       "attrs": dict(comp)}
            out[g] = (dims, planes[k] if k < len(planes) else [0.0] * n)
    return out


def particle_record(tree, i, key, units, np_, columns=None, value=None, dtype="<f8",
                    withcomponents=False, weighted=False, unit_si=1.0, offset=0.0):
    """save_record(it, key, ::ParticleRecord) (hdf5.jl:69-84) of a record
    built by init!(ParticleRecord) (particles.jl:27-51).  A single-value
    record (value given) is a group with value and shape = [np]; otherwise
    columns are the np-long components.  Returns {path: (dims, values)}."""
    f = f"/data/{i}/particles/{key}"
    comp = {"unitSI": f64(unit_si)}
    rec = {"unitDimension": f64v(UNIT_DIMENSION[units]), "timeOffset": f64(offset),
           "macroWeighted": ["u8", 1 if weighted else 0], "weightingPower": f64(1.0)}
    _ensure_groups(tree, f)
    out = {}
    if value is not None:
        tree[f] = {"kind": "group",
                   "attrs": {"value": f64(value), "shape": ["i64v", [int(np_)]], **comp, **rec}}
        return out
    if not withcomponents:
        tree[f] = {"kind": "dataset", "dtype": dtype, "shape": [int(np_)],
                   "attrs": {**comp, **rec}}
        out[f] = ((int(np_),), columns[0])
    else:
        tree[f] = {"kind": "group", "attrs": rec}
        for k, ax in enumerate("xyz"):
            g = f + "/" + ax
            tree[g] = {"kind": "dataset", "dtype": dtype, "shape": [int(np_)], "attrs": dict(comp)}
            out[g] = ((int(np_),), columns[k] if k < len(columns) else [0.0] * int(np_))
    return out


def species_records(tree, i, name, np_, mass, charge, w0, x_cols, v_cols, ids):
    """The nine records the loop registers per kinetic species
    (ParticleInCell.jl:63-71): columns as lists."""
    out = {}
    out.update(particle_record(tree, i, f"{name}/id", "1", np_, [ids], dtype="<u4"))
    particle_record(tree, i, f"{name}/mass", "kg", np_, value=mass)
    particle_record(tree, i, f"{name}/charge", "C", np_, value=charge)
    out.update(particle_record(tree, i, f"{name}/weighting", "1", np_, [[w0] * np_],
                               weighted=True))
    out.update(particle_record(tree, i, f"{name}/momentum", "kg*m/s", np_,
                               [[mass * v for v in col] for col in v_cols], withcomponents=True))
    out.update(particle_record(tree, i, f"{name}/position", "m", np_, x_cols,
                               withcomponents=True))
    for ax in "xyz":
        particle_record(tree, i, f"{name}/positionOffset/{ax}", "m", np_, value=0.0)
    return out
