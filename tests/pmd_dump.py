"""Dump an HDF5 file's tree for the openPMD tests (run by a Python that has
h5py: python pmd_dump.py FILE OUT.json OUT.npz).

Every group and dataset is listed with its attributes, each described by
its HDF5 type and dataspace:
  ["str", v]    fixed-length string, scalar dataspace
  ["null_str"]  fixed-length string, null dataspace
  ["vstr", [..]]  1-D array of variable-length strings
  ["chars", [..]] 1-D array of one-byte fixed-length strings
  ["f64", v] / ["f64v", [..]]   64-bit float scalar / 1-D array
  ["i64", v] / ["i64v", [..]]   64-bit signed integer scalar / 1-D array
  ["u32", v], ["u8", v]         unsigned integer scalars
Every string attribute must be UTF-8 (checked here).  Dataset values go to
OUT.npz under their path with '/' replaced by '|'.
"""
import json
import sys

import h5py
import numpy as np


def _s(b):
    return b.decode("utf-8") if isinstance(b, bytes) else str(b)


def describe(obj, name):
    aid = h5py.h5a.open(obj.id, name.encode())
    t, sp = aid.get_type(), aid.get_space()
    cls, stype = t.get_class(), sp.get_simple_extent_type()
    rank = sp.get_simple_extent_ndims() if stype == h5py.h5s.SIMPLE else 0
    if cls == h5py.h5t.STRING:
        if t.get_cset() != h5py.h5t.CSET_UTF8:
            raise SystemExit(f"{obj.name}@{name}: string not UTF-8")
        if stype == h5py.h5s.NULL:
            return ["null_str"]
        if t.is_variable_str():
            if rank != 1:
                raise SystemExit(f"{obj.name}@{name}: variable string of rank {rank}")
            return ["vstr", [_s(v) for v in obj.attrs[name]]]
        if stype == h5py.h5s.SCALAR:
            if t.get_strpad() != h5py.h5t.STR_NULLTERM:
                raise SystemExit(f"{obj.name}@{name}: string not null-terminated")
            return ["str", _s(obj.attrs[name]), int(t.get_size())]
        if rank == 1 and t.get_size() == 1:
            return ["chars", [_s(v) for v in obj.attrs[name]]]
        raise SystemExit(f"{obj.name}@{name}: unexpected fixed string array")
    v = obj.attrs[name]
    if cls == h5py.h5t.FLOAT and t.get_size() == 8:
        return ["f64", float(v)] if rank == 0 else ["f64v", [float(x) for x in v]]
    if cls == h5py.h5t.INTEGER:
        signed, size = t.get_sign() == h5py.h5t.SGN_2, t.get_size()
        if signed and size == 8:
            return ["i64", int(v)] if rank == 0 else ["i64v", [int(x) for x in v]]
        if not signed and size == 4 and rank == 0:
            return ["u32", int(v)]
        if not signed and size == 1 and rank == 0:
            return ["u8", int(v)]
    raise SystemExit(f"{obj.name}@{name}: unexpected attribute type class {cls} size "
                     f"{t.get_size()} rank {rank}")


def main(path, out_json, out_npz):
    tree, data = {}, {}
    with h5py.File(path, "r") as f:
        def visit(name, obj):
            key = "/" + name
            ent = {"attrs": {a: describe(obj, a) for a in obj.attrs}}
            if isinstance(obj, h5py.Dataset):
                ent["kind"] = "dataset"
                ent["dtype"] = obj.dtype.str
                ent["shape"] = list(obj.shape)
                ent["layout"] = "contiguous" if obj.chunks is None else "chunked"
                data[key.replace("/", "|")] = obj[()]
            else:
                ent["kind"] = "group"
            tree[key] = ent
        tree["/"] = {"kind": "group", "attrs": {a: describe(f, a) for a in f.attrs}}
        f.visititems(visit)
    with open(out_json, "w") as fh:
        json.dump(tree, fh, indent=1, sort_keys=True)
    np.savez(out_npz, **data)


if __name__ == "__main__":
    main(*sys.argv[1:4])
