"""openPMD writer (libiskra_pmd.so) on the CPU: units against their SI
definitions, and files written through the reference-named registry
(iskra_amd.diagnostics) read back with h5py and compared, object by object
and attribute by attribute, with the layout the reference's
Diagnostics/src/hdf5.jl writes (restated in oracle/openpmd_layout.py)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from iskra_amd import _lib as L
from iskra_amd import diagnostics as diag
from iskra_amd import particle_in_cell as pic
from iskra_amd import regular_grids as rg
from oracle import openpmd_layout as OL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H5PY_PYTHON = "/opt/conda/bin/python3.9"
H5DUMP = "/opt/conda/bin/h5dump"
DATE = re.compile(r"^\d{4}/\d{1,2}/\d{1,2} \d{2}:\d{2}$")


def read_back(path, tmp_path):
    """(tree, datasets) of an HDF5 file, read by h5py in a separate Python."""
    if not os.path.exists(H5PY_PYTHON):
        pytest.fail(f"{H5PY_PYTHON} (the h5py reader) is missing")
    js, nz = str(tmp_path / "dump.json"), str(tmp_path / "dump.npz")
    env = {k: v for k, v in os.environ.items() if not k.startswith("PYTHON")}
    env["PYTHONNOUSERSITE"] = "1"
    r = subprocess.run([H5PY_PYTHON, os.path.join(ROOT, "tests", "pmd_dump.py"), path, js, nz],
                       capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(js) as fh:
        tree = json.load(fh)
    with np.load(nz) as z:
        data = {k.replace("|", "/"): z[k] for k in z.files}
    return tree, data


def compare(tree, data, expected, datasets):
    """Every object and attribute of the file equals the expectation (the
    root's date only by format); dataset values are compared exactly."""
    date = tree["/"]["attrs"].pop("date")
    assert date[0] == "str" and DATE.match(date[1]), date
    assert sorted(tree) == sorted(expected), (set(tree) ^ set(expected))
    for path, ent in expected.items():
        got = tree[path]
        assert got["kind"] == ent["kind"], path
        assert got["attrs"] == ent["attrs"], (path, got["attrs"], ent["attrs"])
        if ent["kind"] == "dataset":
            assert got["dtype"] == ent["dtype"], path
            assert got["shape"] == ent["shape"], path
            assert got["layout"] == "contiguous", path
    for path, (dims, values) in datasets.items():
        a = data[path]
        assert a.shape == tuple(dims), path
        assert np.array_equal(a.ravel(), np.asarray(values, dtype=a.dtype).ravel()), path


def test_units_against_si_definitions():
    # every unit string of the reference's loop
    for u, dims in OL.UNIT_DIMENSION.items():
        d, si = diag.units(u)
        assert d == tuple(float(x) for x in dims), u
        assert si == 1.0, u
    # composition: a product adds dimensions, a quotient subtracts them
    names = list(OL.UNIT_DIMENSION)
    for a in names:
        for b in names:
            da, db = diag.units(a)[0], diag.units(b)[0]
            assert diag.units(f"({a})*({b})")[0] == tuple(x + y for x, y in zip(da, db))
            assert diag.units(f"({a})/({b})")[0] == tuple(x - y for x, y in zip(da, db))
    # prefixes and non-coherent units: exact SI values
    assert diag.units("cm^-3") == ((-3.0, 0, 0, 0, 0, 0, 0), 1e6)
    assert diag.units("eV") == ((2.0, 1, -2, 0, 0, 0, 0), 1.602176634e-19)
    assert diag.units("keV")[1] == 1.602176634e-19 * 1e3
    assert diag.units("mm/ms") == diag.units("m/s")
    assert diag.units("kg*m^2/s^2") == diag.units("J") == diag.units("N*m")
    assert diag.units("Ohm") == diag.units("V/A")
    assert diag.units("m^-2") == diag.units("m^(-2)") == diag.units("1/m^2")
    assert diag.units("m^(1/2)")[0][0] == 0.5
    assert diag.units("Torr")[1] == 101325.0 / 760.0
    for bad in ("", "furlong", "m^", "m)", "(m", "2*", "m**2", "-1/m", "kmin"):
        with pytest.raises(L.IskraError) as e:
            diag.units(bad)
        assert e.value.code == L.E_ARG, bad


def _species(name, n, D, q, m, w0, seed):
    part = pic.create_kinetic_species(name, 2 * n, q, m, w0, D=D)
    rng = np.random.default_rng(seed)
    part.x[:n] = rng.random((n, D))
    part.v[:n] = rng.normal(size=(n, 3)) * 1e5
    part.id[:n] = rng.permutation(np.arange(1, n + 1, dtype=np.uint32))
    part.np = n
    return part


def _register_species(part):
    """the reference's nine @particle calls (ParticleInCell.jl:63-71)"""
    nm = part.name
    diag.particle(nm + "/id", "1", part.id, part)
    diag.particle(nm + "/mass", "kg", [part.m], part)
    diag.particle(nm + "/charge", "C", [part.q], part)
    diag.particle(nm + "/weighting", "1", part.wg, part, weighted=True)
    diag.particle(nm + "/momentum", "kg*m/s", part.m * part.v, part, withcomponents=True)
    diag.particle(nm + "/position", "m", part.x, part, withcomponents=True)
    for ax in "xyz":
        diag.particle(f"{nm}/positionOffset/{ax}", "m", [0.0], part)


def _expect_species(tree, i, part):
    n = part.np
    return OL.species_records(tree, i, part.name, n, part.m, part.q, part.w0,
                              [list(part.x[:n, k]) for k in range(part.D)],
                              [list(part.v[:n, k]) for k in range(3)], list(part.id[:n]))


def test_file_layout_2d(tmp_path):
    diag.records.clear()
    nx, ny = 9, 6
    grid = rg.create_uniform_grid(np.linspace(0.0, 0.08, nx), np.linspace(0.0, 0.05, ny))
    rng = np.random.default_rng(1)
    rho = rng.normal(size=(nx, ny))
    E = rng.normal(size=(nx, ny, 3))
    diag.field("rho", "C/m^2", rho, grid)
    diag.field("E", "V/m", E, grid, withcomponents=True)
    diag.field("Exy", "V/m", E[..., :2], grid, withcomponents=True)
    diag.probe("I1", 0.25, "A")
    e = _species("e-", 37, 2, -1.602176634e-19, 9.1093837015e-31, 1e3, 2)
    _register_species(e)
    path = diag.new_iteration(lambda it: (diag.save_records(it, "e-/"),
                                          diag.save_record(it, "rho"),
                                          diag.save_record(it, "E"),
                                          diag.save_record(it, "Exy"),
                                          diag.save_record(it, "I1")),
                              str(tmp_path / "run"), 40, 3.9e-9, 1e-10, seed=7)
    assert path == str(tmp_path / "run" / "hdf5" / "data40.h5")
    tree, data = read_back(path, tmp_path)

    exp = OL.skeleton(40, 3.9e-9, 1e-10, seed=7)
    g = dict(dim=2, axial=0, nodes=(nx, ny), spacing=(0.01, 0.01))
    ds = {}
    ds.update(OL.field_record(exp, 40, "rho", "C/m^2", g, [list(rho.ravel(order="F"))], False))
    ds.update(OL.field_record(exp, 40, "E", "V/m", g,
                              [list(E[..., k].ravel(order="F")) for k in range(3)], True))
    ds.update(OL.field_record(exp, 40, "Exy", "V/m", g,
                              [list(E[..., k].ravel(order="F")) for k in range(2)], True))
    ds.update(OL.field_record(exp, 40, "I1", "A", None, [[0.25]], False, probe=True))
    ds.update(_expect_species(exp, 40, e))
    compare(tree, data, exp, ds)
    # orientation: element [j, i] of a field dataset is the node (x_i, y_j)
    assert np.array_equal(data["/data/40/fields/rho"], rho.T)
    assert np.array_equal(data["/data/40/fields/E/y"], E[..., 1].T)


def test_file_layout_1d_and_axial(tmp_path):
    diag.records.clear()
    g1 = rg.create_uniform_grid(np.linspace(0.0, 1.0, 17))
    phi = np.linspace(-1.0, 1.0, 17) ** 3
    E1 = np.stack([np.linspace(0, 2, 17), np.zeros(17), np.zeros(17)], axis=1)
    diag.field("phi", "V", phi, g1)
    diag.field("E", "V/m", E1, g1, withcomponents=True)
    ga = rg.create_axial_grid(np.linspace(0.0, 0.02, 5), np.linspace(0.0, 0.03, 7))
    nu = np.arange(35.0).reshape(5, 7)
    diag.field("nuMCC-e--1", "1/m^2", nu, ga, pos=(0.5, 0.25), offset=2.0)
    ion = _species("He+", 5, 1, 1.602176634e-19, 6.64e-27, 10.0, 3)
    _register_species(ion)
    path = diag.new_iteration(lambda it: (diag.save_record(it, "phi"),
                                          diag.save_record(it, "E"),
                                          diag.save_record(it, "nuMCC-e--1"),
                                          diag.save_records(it, "He+/")),
                              str(tmp_path / "one"), 3, 0.0, 0.5)
    tree, data = read_back(path, tmp_path)
    exp = OL.skeleton(3, 0.0, 0.5)
    gd1 = dict(dim=1, axial=0, nodes=(17,), spacing=(1.0 / 16,))
    gda = dict(dim=2, axial=1, nodes=(5, 7), spacing=(0.005, 0.005))
    ds = {}
    ds.update(OL.field_record(exp, 3, "phi", "V", gd1, [list(phi)], False))
    ds.update(OL.field_record(exp, 3, "E", "V/m", gd1, [list(E1[:, k]) for k in range(3)], True))
    ds.update(OL.field_record(exp, 3, "nuMCC-e--1", "1/m^2", gda, [list(nu.ravel(order="F"))],
                              False, position=(0.5, 0.25), offset=2.0))
    ds.update(_expect_species(exp, 3, ion))
    compare(tree, data, exp, ds)
    assert data["/data/3/fields/phi"].shape == (1, 17)


def test_empty_species_records(tmp_path):
    """a species with no particles: zero-length datasets, shape [0]"""
    diag.records.clear()
    part = _species("e-", 4, 2, -1.0, 1.0, 1.0, 5)
    part.np = 0
    _register_species(part)
    path = diag.new_iteration(lambda it: diag.save_records(it, "e-/"), str(tmp_path), 2, 0.0, 1.0)
    tree, data = read_back(path, tmp_path)
    exp = OL.skeleton(2, 0.0, 1.0)
    ds = OL.species_records(exp, 2, "e-", 0, part.m, part.q, part.w0, [[], []],
                            [[], [], []], [])
    compare(tree, data, exp, ds)


def test_h5dump_reads_the_file(tmp_path):
    diag.records.clear()
    grid = rg.create_uniform_grid(np.linspace(0.0, 1.0, 5), np.linspace(0.0, 1.0, 4))
    diag.field("rho", "C/m^2", np.ones((5, 4)), grid)
    path = diag.new_iteration(lambda it: diag.save_record(it, "rho"), str(tmp_path), 0, 0.0, 1.0)
    if not os.path.exists(H5DUMP):
        pytest.fail(f"{H5DUMP} is missing")
    r = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for word in ('GROUP "data"', 'GROUP "0"', 'GROUP "fields"', 'GROUP "particles"',
                 'DATASET "rho"', "DATASPACE  SIMPLE { ( 4, 5 ) / ( 4, 5 ) }",
                 'ATTRIBUTE "axisLabels"', 'ATTRIBUTE "openPMD"', "H5T_CSET_UTF8"):
        assert word in r.stdout, word


def test_writer_errors(tmp_path, capsys):
    diag.records.clear()
    grid = rg.create_uniform_grid(np.linspace(0.0, 1.0, 5))
    diag.field("rho", "C/m^2", np.ones(5), grid)
    diag.field("bad", "parsec", np.ones(5), grid)
    part = _species("e-", 3, 1, -1.0, 1.0, 1.0, 4)
    diag.particle("e-/id", "1", part.id, part)
    diag.records["noslash"] = diag.records["e-/id"]
    kept = []

    def body(it):
        kept.append(it)
        diag.save_record(it, "rho")
        with pytest.raises(L.IskraError) as e:
            diag.save_record(it, "rho")
        assert e.value.code == L.E_STATE
        with pytest.raises(L.IskraError) as e:
            diag.save_record(it, "bad")
        assert e.value.code == L.E_ARG and "parsec" in str(e.value)
        with pytest.raises(L.IskraError) as e:
            diag.save_record(it, "noslash")
        assert e.value.code == L.E_ARG
        diag.save_record(it, "missing")
        raise KeyError("raised inside the block")

    with pytest.raises(KeyError):
        diag.new_iteration(body, str(tmp_path), 1, 0.0, 1.0)
    assert "Couldn't find diagnostic missing" in capsys.readouterr().out
    with pytest.raises(L.IskraError) as e:   # the block's file was closed
        diag.save_record(kept[0], "e-/id")
    assert e.value.code == L.E_STATE
    tree, _ = read_back(str(tmp_path / "hdf5" / "data1.h5"), tmp_path)
    assert "/data/1/fields/rho" in tree and "/data/1/fields/bad" not in tree
    with pytest.raises(ValueError):
        diag.field("rho", "C/m^2", np.ones(4), grid)


def test_c_abi_rejects_bad_arguments(tmp_path):
    """NULL pointers, bad grids, component counts and dtypes are errors,
    not crashes, at the C boundary (include/iskra_pmd.h)"""
    import ctypes as C
    lib = diag.pmd_lib()
    h = C.c_void_p()
    assert lib.iskra_pmd_open(None, 0, 0.0, 1.0, None, C.byref(h)) == L.E_ARG
    assert lib.iskra_pmd_open(str(tmp_path).encode(), -1, 0.0, 1.0, None, C.byref(h)) == L.E_ARG
    assert lib.iskra_pmd_open(str(tmp_path).encode(), 5, 0.0, 1.0, None, C.byref(h)) == L.OK
    data = np.zeros(9)
    ptr = data.ctypes.data_as(C.c_void_p)
    g = diag.Grid(dim=3, axial=0)
    g.nodes[0], g.nodes[1] = 3, 3
    assert lib.iskra_pmd_field(h, b"rho", b"C", C.byref(g), 1, 0, ptr, None, 0.0) == L.E_ARG
    g.dim = 2
    assert lib.iskra_pmd_field(h, b"rho", b"C", None, 1, 0, ptr, None, 0.0) == L.E_ARG
    assert lib.iskra_pmd_field(h, b"rho", b"C", C.byref(g), 2, 0, ptr, None, 0.0) == L.E_ARG
    assert lib.iskra_pmd_field(h, b"/rho", b"C", C.byref(g), 1, 0, ptr, None, 0.0) == L.E_ARG
    assert lib.iskra_pmd_field(h, b"rho", b"C", C.byref(g), 1, 0, ptr, None, 0.0) == L.OK
    assert lib.iskra_pmd_particle(h, b"e-/x", b"m", 4, 1, 0, 7, ptr, 4, 0, 0.0) == L.E_ARG
    assert lib.iskra_pmd_particle(h, b"e-/x", b"m", 4, 1, 0, 0, ptr, 3, 0, 0.0) == L.E_ARG
    assert lib.iskra_pmd_particle(h, b"e-/x", b"m", 4, 1, 0, 0, None, 4, 0, 0.0) == L.E_ARG
    assert lib.iskra_pmd_particle_constant(h, b"e-/m", b"kg", 1.0, -1, 0, 0.0) == L.E_ARG
    assert lib.iskra_pmd_probe(h, b"", b"V", 1.0, 0.0) == L.E_ARG
    assert lib.iskra_pmd_close(h) == L.OK
    assert lib.iskra_pmd_close(None) == L.E_ARG
    assert lib.iskra_pmd_field(None, b"rho", b"C", C.byref(g), 1, 0, ptr, None, 0.0) == L.E_STATE


def test_pmd_library_exports_the_header():
    hdr = open(os.path.join(ROOT, "include", "iskra_pmd.h")).read()
    names = set(re.findall(r"\b(iskra_pmd_\w+)\s*\(", hdr))
    assert names == set(diag.PMD_EXPORTS)
    lib = diag.pmd_lib()
    for n in names:
        assert hasattr(lib, n), n
