"""ctypes binding of libiskra_hip.so (the C ABI declared in include/iskra.h).

The library is built in-tree by `python -m iskra_amd.build`; importing this
module never falls back to anything else: if the shared object is missing or
no GPU is visible, calls raise IskraError.
"""
import atexit
import ctypes as C
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libiskra_hip.so")

OK, E_ARG, E_HIP, E_CAPACITY, E_PROB, E_STATE, E_CONVERGE, E_UNSUPPORTED, E_IO = range(9)
BND_NONE, BND_WRAP, BND_DISCARD = 0, 1, 2
F_RHO, F_PHI, F_EX, F_EY, F_BG, F_DENSITY, F_TARGET, F_FLUID_N, F_FLUID_DN = range(9)
POISSON_DIRECT, POISSON_PCG, POISSON_MG, POISSON_MGPCG, POISSON_SPECTRAL = range(5)
PUSHER_XY, PUSHER_RZ = 0, 1
SURF_ABSORBING, SURF_REFLECTIVE, SURF_PERIODIC = 1, 2, 3
WARM_ZERO, WARM_PREV = 0, 1
WS_PHI_MSE, WS_ENERGY, WS_ENERGY_NET = 0, 1, 2
COLL_ELASTIC_ISOTROPIC, COLL_ELASTIC_BACKWARD, COLL_EXCITATION, COLL_IONIZATION = range(4)
COLL_INELASTIC_BACKWARD = 4
NOALLOC_SCRATCH_BYTES = 4 << 20

EXPORTS = [
    "iskra_last_error", "iskra_version", "iskra_ctx_create", "iskra_ctx_destroy", "iskra_sync",
    "iskra_grid_set", "iskra_grid_set_axial", "iskra_pusher_set", "iskra_dsmc_add",
    "iskra_dsmc_run", "iskra_dsmc_counts", "iskra_tracker_create", "iskra_tracker_surface",
    "iskra_tracker_table", "iskra_electrode_create", "iskra_electrode_state",
    "iskra_circuit_set", "iskra_circuit_state", "iskra_fluid_add", "iskra_fluid_advance", "iskra_checkpoint_size", "iskra_checkpoint_save", "iskra_checkpoint_load", "iskra_tuning_set", "iskra_tuning_get", "iskra_species_add",
    "iskra_species_set_boundary", "iskra_species_sort_every", "iskra_species_margin",
    "iskra_species_upload",
    "iskra_species_download",
    "iskra_species_count", "iskra_species_get_state", "iskra_species_set_state",
    "iskra_species_sample_maxwellian", "iskra_sort", "iskra_gather",
    "iskra_push", "iskra_deposit_species", "iskra_deposit", "iskra_field_download",
    "iskra_field_upload", "iskra_poisson_set", "iskra_poisson_set_tag_value",
    "iskra_poisson_set_tag_waveform", "iskra_poisson_solve", "iskra_efield", "iskra_mcc_add",
    "iskra_mcc_run", "iskra_mcc_counts", "iskra_mcc_counts_last", "iskra_dsmc_counts_last", "iskra_rng_set", "iskra_rng_export",
    "iskra_rng_export_noalloc", "iskra_step",
    "iskra_stats_get", "iskra_profile", "iskra_bench_copy",
    "iskra_ws_create", "iskra_ws_destroy", "iskra_ws_param_count", "iskra_ws_set_params",
    "iskra_ws_get_params", "iskra_ws_set_gain", "iskra_ws_forward", "iskra_ws_backward",
    "iskra_ws_dataset_reserve", "iskra_ws_capture", "iskra_ws_dataset_size",
    "iskra_ws_dataset_rms", "iskra_ws_train", "iskra_ws_eval", "iskra_ws_predict",
    "iskra_ws_set_objective", "iskra_ws_solve", "iskra_ws_capture_iterates",
    "iskra_ws_dataset_get",
    "iskra_batch_create", "iskra_batch_destroy", "iskra_batch_set_graph", "iskra_batch_step",
    "iskra_batch_graphs_launched",
]


class IskraError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"iskra error {code}: {msg}")
        self.code = code


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_int32), ("chunk", C.c_int32), ("margin", C.c_int32),
                ("sort_every", C.c_int32)]


K_PUSH, K_SORT, K_POISSON, K_MCC, K_MERGE, K_EFIELD, K_WARMSTART = range(7)
K_NAMES = ["push", "sort", "poisson", "mcc", "merge", "efield", "warmstart", "k7"]


class Stats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("particle_steps", C.c_int64),
                ("poisson_solves", C.c_int64), ("poisson_iters", C.c_int64),
                ("last_iters", C.c_int32), ("pad", C.c_int32), ("last_relres", C.c_double),
                ("kernel_ms", C.c_double * 8), ("kernel_calls", C.c_int64 * 8),
                ("push_particles", C.c_int64)]


_lib = None
P = C.c_void_p


def lib():
    """Load libiskra_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise IskraError(E_STATE, f"{LIBPATH} is missing: run `python -m iskra_amd.build`")
        L = C.CDLL(LIBPATH)
        I, I64, D, U32, U64 = C.c_int, C.c_int64, C.c_double, C.c_uint32, C.c_uint64
        sig = {
            "iskra_last_error": (C.c_char_p, []),
            "iskra_version": (I, []),
            "iskra_ctx_create": (I, [I, C.POINTER(P)]),
            "iskra_ctx_destroy": (I, [P]),
            "iskra_sync": (I, [P]),
            "iskra_grid_set": (I, [P, I, P, P, P, P]),
            "iskra_grid_set_axial": (I, [P, P, P, P, P]),
            "iskra_pusher_set": (I, [P, I]),
            "iskra_dsmc_add": (I, [P, I, I, I, P, P, C.POINTER(I)]),
            "iskra_dsmc_run": (I, [P, I, D, C.POINTER(I64)]),
            "iskra_dsmc_counts": (I, [P, I, I, P]),
            "iskra_tracker_create": (I, [P, I]),
            "iskra_tracker_surface": (I, [P, P, I]),
            "iskra_tracker_table": (I, [P, P, P]),
            "iskra_electrode_create": (I, [P, P, I, D, D, C.POINTER(I)]),
            "iskra_electrode_state": (I, [P, I, P, P, P, P]),
            "iskra_circuit_set": (I, [P, D, D, D, D, D, D, D, I, I]),
            "iskra_circuit_state": (I, [P, P, P, P, P]),
            "iskra_fluid_add": (I, [P, D, D, D, P, C.POINTER(I)]),
            "iskra_fluid_advance": (I, [P, D]),
            "iskra_checkpoint_size": (I, [P, C.POINTER(I64)]),
            "iskra_checkpoint_save": (I, [P, P, I64]),
            "iskra_checkpoint_load": (I, [P, P, I64]),
            "iskra_tuning_set": (I, [P, C.POINTER(Tuning)]),
            "iskra_tuning_get": (I, [P, C.POINTER(Tuning)]),
            "iskra_species_add": (I, [P, D, D, D, I64, I, C.POINTER(I)]),
            "iskra_species_set_boundary": (I, [P, I, P]),
            "iskra_species_sort_every": (I, [P, I, I]),
            "iskra_species_margin": (I, [P, I, I]),
            "iskra_species_upload": (I, [P, I, I64, P, P, P]),
            "iskra_species_download": (I, [P, I, I64, C.POINTER(I64), P, P, P]),
            "iskra_species_count": (I, [P, I, C.POINTER(I64)]),
            "iskra_species_get_state": (I, [P, I, I64, C.POINTER(I64), P, P, P, P, P, P, P]),
            "iskra_species_set_state": (I, [P, I, I64, P, P, P, P, P, P, P]),
            "iskra_species_sample_maxwellian": (I, [P, I, I64, P, P, P, P, U64, U32]),
            "iskra_sort": (I, [P, I, I]),
            "iskra_gather": (I, [P, I, I64, P]),
            "iskra_push": (I, [P, I, D]),
            "iskra_deposit_species": (I, [P, I]),
            "iskra_deposit": (I, [P]),
            "iskra_field_download": (I, [P, I, I, P]),
            "iskra_field_upload": (I, [P, I, I, P]),
            "iskra_poisson_set": (I, [P, I, D, P, I, I, D, I]),
            "iskra_poisson_set_tag_value": (I, [P, I, D]),
            "iskra_poisson_set_tag_waveform": (I, [P, I, D, D, D, D]),
            "iskra_poisson_solve": (I, [P, I, C.POINTER(I), C.POINTER(D)]),
            "iskra_efield": (I, [P]),
            "iskra_mcc_add": (I, [P, I, I, P, P, P, P, P, I, D, D, D, C.POINTER(I)]),
            "iskra_mcc_run": (I, [P, I, D, P]),
            "iskra_mcc_counts": (I, [P, I, I, P]),
            "iskra_mcc_counts_last": (I, [P, I, P]),
            "iskra_dsmc_counts_last": (I, [P, I, P]),
            "iskra_rng_set": (I, [P, U64]),
            "iskra_rng_export": (I, [P, U32, U32, P, I64, I, P]),
            "iskra_rng_export_noalloc": (I, [P, U32, U32, P, I64, I, P]),
            "iskra_step": (I, [P, D, I64]),
            "iskra_stats_get": (I, [P, C.POINTER(Stats)]),
            "iskra_profile": (I, [P, I]),
            "iskra_bench_copy": (I, [P, I64, I, I, C.POINTER(D)]),
            "iskra_ws_create": (I, [P, I, I, I, I, C.POINTER(P)]),
            "iskra_ws_destroy": (I, [P]),
            "iskra_ws_param_count": (I, [P, C.POINTER(I64)]),
            "iskra_ws_set_params": (I, [P, P]),
            "iskra_ws_get_params": (I, [P, P]),
            "iskra_ws_set_gain": (I, [P, D]),
            "iskra_ws_forward": (I, [P, I, P, P]),
            "iskra_ws_backward": (I, [P, I, P, P, P, C.POINTER(D)]),
            "iskra_ws_dataset_reserve": (I, [P, I64]),
            "iskra_ws_capture": (I, [P, C.POINTER(I64)]),
            "iskra_ws_dataset_size": (I, [P, C.POINTER(I64)]),
            "iskra_ws_dataset_rms": (I, [P, I64, I64, C.POINTER(D)]),
            "iskra_ws_train": (I, [P, I64, I64, I, I, C.c_float, U64, P]),
            "iskra_ws_eval": (I, [P, I64, I64, C.POINTER(D)]),
            "iskra_ws_predict": (I, [P]),
            "iskra_ws_set_objective": (I, [P, I, D, I]),
            "iskra_ws_solve": (I, [P, I, D, I, I, C.POINTER(I), C.POINTER(D)]),
            "iskra_ws_capture_iterates": (I, [P, D, I, P, I, C.POINTER(I64)]),
            "iskra_ws_dataset_get": (I, [P, I64, P, P]),
            "iskra_batch_create": (I, [P, I, C.POINTER(P)]),
            "iskra_batch_destroy": (I, [P]),
            "iskra_batch_set_graph": (I, [P, I]),
            "iskra_batch_step": (I, [P, D, I64]),
            "iskra_batch_graphs_launched": (I, [P, C.POINTER(I64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().iskra_last_error().decode(errors="replace")
        raise IskraError(rc, msg)


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P)


_live_ws = weakref.WeakSet()
_live_batch = weakref.WeakSet()
_live_ctx = weakref.WeakSet()


def close_all():
    """Destroy every live warm-start network, then every live context, in that
    order (networks use their context's stream).  Registered with atexit so
    that teardown happens while the interpreter and the library are intact."""
    for w in list(_live_ws) + list(_live_batch):
        try:
            w.close()
        except Exception:
            pass
    for c in list(_live_ctx):
        try:
            c.close()
        except Exception:
            pass


atexit.register(close_all)


class Context:
    """One device context: grid, species, fields, Poisson solver, MCC."""

    def __init__(self, device=0):
        h = P()
        check(lib().iskra_ctx_create(device, C.byref(h)))
        self.h = h
        self.dim = None
        self.nodes = None
        _live_ctx.add(self)

    def close(self):
        if self.h:
            for w in list(_live_ws):
                if w.ctx is self:
                    w.close()
            for bt in list(_live_batch):
                if any(c is self for c in bt.ctxs):
                    bt.close()
            lib().iskra_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # grid
    def grid(self, dim, nodes, dh, origin=(0.0, 0.0), vol_periodic=(0, 0, 0, 0)):
        n = np.asarray(list(nodes) + [1] * (2 - len(nodes)), dtype=np.int64)
        d = np.asarray(list(dh) + [1.0] * (2 - len(dh)), dtype=np.float64)
        o = np.asarray(list(origin) + [0.0] * (2 - len(origin)), dtype=np.float64)
        vp = np.asarray(vol_periodic, dtype=np.int32)
        check(lib().iskra_grid_set(self.h, dim, _p(n), _p(d), _p(o), _p(vp)))
        self.dim = dim
        self.nodes = (int(n[0]), int(n[1]) if dim == 2 else 1)

    def grid_axial(self, nodes, dh, origin=(0.0, 0.0), vol_periodic_z=(0, 0)):
        """(r, z) grid: nodes (nr, nz), dh (dr, dz); selects the RZ pusher."""
        n = np.asarray(nodes, dtype=np.int64)
        d = np.asarray(dh, dtype=np.float64)
        o = np.asarray(origin, dtype=np.float64)
        vp = np.asarray(vol_periodic_z, dtype=np.int32)
        check(lib().iskra_grid_set_axial(self.h, _p(n), _p(d), _p(o), _p(vp)))
        self.dim = 2
        self.nodes = (int(n[0]), int(n[1]))

    def pusher(self, kind):
        check(lib().iskra_pusher_set(self.h, int(kind)))

    @property
    def nn(self):
        return self.nodes[0] * self.nodes[1]

    def tuning(self, tile=0, chunk=0, margin=0, sort_every=-1):
        t = Tuning(tile, chunk, margin, sort_every)
        check(lib().iskra_tuning_set(self.h, C.byref(t)))

    def get_tuning(self):
        t = Tuning()
        check(lib().iskra_tuning_get(self.h, C.byref(t)))
        return dict(tile=t.tile, chunk=t.chunk, margin=t.margin, sort_every=t.sort_every)

    # species
    def species(self, q, m, w0, capacity, track_ids=True):
        s = C.c_int()
        check(lib().iskra_species_add(self.h, q, m, w0, int(capacity), int(track_ids),
                                      C.byref(s)))
        return s.value

    def species_sort_every(self, sid, every):
        """Steps between this species' sorts in step() (0: the context's)."""
        check(lib().iskra_species_sort_every(self.h, sid, int(every)))

    def species_margin(self, sid, margin):
        """Deposit-window margin of one species in cells (-1: the context's)."""
        check(lib().iskra_species_margin(self.h, sid, int(margin)))

    def boundary(self, sid, modes):
        md = np.asarray(list(modes) + [BND_WRAP] * (2 - len(modes)), dtype=np.int32)
        check(lib().iskra_species_set_boundary(self.h, sid, _p(md)))

    def upload(self, sid, x, v, ids=None):
        """x: (np, D) FP64, v: (np, 3) FP64 (converted to column-major)."""
        x = np.asarray(x, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        n = x.shape[0]
        xc = np.ascontiguousarray(x.T).ravel()
        vc = np.ascontiguousarray(v.T).ravel()
        idc = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint32)
        check(lib().iskra_species_upload(self.h, sid, n, _p(xc), _p(vc), _p(idc)))

    def download(self, sid, cap=None):
        if cap is None:
            cap = self.count_raw(sid)
        cap = max(int(cap), 1)
        x = np.zeros(cap * self.dim)
        v = np.zeros(cap * 3)
        ids = np.zeros(cap, dtype=np.uint32)
        n = C.c_int64()
        check(lib().iskra_species_download(self.h, sid, cap, C.byref(n), _p(x), _p(v), _p(ids)))
        n = n.value
        X = x.reshape(self.dim, cap)[:, :n].T.copy()
        V = v.reshape(3, cap)[:, :n].T.copy()
        return X, V, ids[:n].copy()

    def get_state(self, sid):
        """Raw device state (current order, discarded particles included)."""
        n = C.c_int64()
        check(lib().iskra_species_get_state(self.h, sid, 0, C.byref(n), None, None, None, None,
                                            None, None, None))
        n = n.value
        cap = max(n, 1)
        st = dict(cell=np.zeros(cap, np.int32), ox=np.zeros(cap, np.float32),
                  oy=np.zeros(cap, np.float32), vx=np.zeros(cap, np.float32),
                  vy=np.zeros(cap, np.float32), vz=np.zeros(cap, np.float32),
                  id=np.zeros(cap, np.uint32))
        m = C.c_int64()
        check(lib().iskra_species_get_state(self.h, sid, cap, C.byref(m), _p(st["cell"]),
                                            _p(st["ox"]), _p(st["oy"]), _p(st["vx"]),
                                            _p(st["vy"]), _p(st["vz"]), _p(st["id"])))
        return {k: v[:n].copy() for k, v in st.items()}

    def set_state(self, sid, cell, ox, oy, vx, vy, vz, ids=None):
        n = len(cell)
        a = [np.ascontiguousarray(cell, np.int32), np.ascontiguousarray(ox, np.float32),
             np.ascontiguousarray(oy if oy is not None else np.zeros(n), np.float32),
             np.ascontiguousarray(vx, np.float32), np.ascontiguousarray(vy, np.float32),
             np.ascontiguousarray(vz, np.float32)]
        idc = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        check(lib().iskra_species_set_state(self.h, sid, n, *[_p(x) for x in a], _p(idc)))

    def count(self, sid):
        n = C.c_int64()
        check(lib().iskra_species_count(self.h, sid, C.byref(n)))
        return n.value

    def count_raw(self, sid):
        return self.count(sid)

    def sample_maxwellian(self, sid, n, wx, wv, dx=None, dv=None, seed=0, stream=0):
        wx = np.asarray(wx, dtype=np.float64)
        wv = np.asarray(wv, dtype=np.float64)
        dx = None if dx is None else np.asarray(dx, dtype=np.float64)
        dv = None if dv is None else np.asarray(dv, dtype=np.float64)
        check(lib().iskra_species_sample_maxwellian(self.h, sid, int(n), _p(wx), _p(dx), _p(wv),
                                                    _p(dv), int(seed), int(stream)))

    # operators
    def sort(self, sid, exact_cell=False):
        check(lib().iskra_sort(self.h, sid, int(exact_cell)))

    def gather(self, sid):
        n = self.count(sid)
        cap = max(n, 1)
        out = np.zeros(3 * cap)
        check(lib().iskra_gather(self.h, sid, cap, _p(out)))
        return out.reshape(3, cap)[:, :n].T.copy()

    def push(self, sid, dt):
        check(lib().iskra_push(self.h, sid, float(dt)))

    def deposit_species(self, sid):
        check(lib().iskra_deposit_species(self.h, sid))

    def deposit(self):
        check(lib().iskra_deposit(self.h))

    # fields (node arrays, flat with i fastest = column-major nx x ny)
    def field(self, which, index=0):
        out = np.zeros(self.nn)
        check(lib().iskra_field_download(self.h, which, index, _p(out)))
        return out

    def set_field(self, which, values, index=0):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        assert v.size == self.nn
        check(lib().iskra_field_upload(self.h, which, index, _p(v)))

    # Poisson
    def poisson(self, kind, eps0, tags=None, periodic_x=False, periodic_y=False, rtol=1e-8,
                max_iter=0):
        t = None if tags is None else np.ascontiguousarray(tags, dtype=np.int8).ravel()
        check(lib().iskra_poisson_set(self.h, kind, float(eps0), _p(t), int(periodic_x),
                                      int(periodic_y), float(rtol), int(max_iter)))

    def tag_value(self, tag, value):
        check(lib().iskra_poisson_set_tag_value(self.h, tag, float(value)))

    def tag_waveform(self, tag, offset, amplitude, freq, phase=0.0):
        check(lib().iskra_poisson_set_tag_waveform(self.h, tag, offset, amplitude, freq, phase))

    def solve(self, warm=WARM_ZERO):
        it = C.c_int()
        rr = C.c_double()
        check(lib().iskra_poisson_solve(self.h, warm, C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def efield(self):
        check(lib().iskra_efield(self.h))

    # MCC
    def mcc(self, src, types, energy, tables, ion=-1, target_m=1.0, target_q=0.0,
            target_T=300.0):
        off = [0]
        e_all, s_all = [], []
        for e, s in tables:
            e_all += list(e)
            s_all += list(s)
            off.append(len(e_all))
        ty = np.asarray(types, dtype=np.int32)
        en = np.asarray(energy, dtype=np.float64)
        off = np.asarray(off, dtype=np.int32)
        e_all = np.asarray(e_all, dtype=np.float64)
        s_all = np.asarray(s_all, dtype=np.float64)
        mid = C.c_int()
        check(lib().iskra_mcc_add(self.h, src, len(types), _p(ty), _p(en), _p(off), _p(e_all),
                                  _p(s_all), ion, target_m, target_q, target_T, C.byref(mid)))
        return mid.value

    def mcc_run(self, mid, dt):
        ev = np.zeros(8, dtype=np.int64)
        check(lib().iskra_mcc_run(self.h, mid, float(dt), _p(ev)))
        return ev

    def mcc_counts(self, mid, ncoll, reset=False):
        out = np.zeros(self.nn * ncoll, dtype=np.int64)
        check(lib().iskra_mcc_counts(self.h, mid, int(reset), _p(out)))
        return out.reshape(ncoll, self.nn)

    def mcc_counts_last(self, mid, ncoll):
        """the counts of the last step of the latest step call"""
        out = np.zeros(self.nn * ncoll, dtype=np.int64)
        check(lib().iskra_mcc_counts_last(self.h, mid, _p(out)))
        return out.reshape(ncoll, self.nn)

    # RNG
    def dsmc(self, src, tgt, tab_g, tab_sigma):
        """One DSMC.ElasticCollision interaction; returns its id."""
        g = np.ascontiguousarray(tab_g, dtype=np.float64)
        s = np.ascontiguousarray(tab_sigma, dtype=np.float64)
        out = C.c_int()
        check(lib().iskra_dsmc_add(self.h, int(src), int(tgt), len(g), _p(g), _p(s),
                                   C.byref(out)))
        return out.value

    def dsmc_run(self, did, dt):
        ev = C.c_int64(0)
        check(lib().iskra_dsmc_run(self.h, int(did), float(dt), C.byref(ev)))
        return ev.value

    def dsmc_counts(self, did, reset=False):
        out = np.zeros(self.nn, dtype=np.int64)
        check(lib().iskra_dsmc_counts(self.h, int(did), int(reset), _p(out)))
        return out

    def dsmc_counts_last(self, did):
        out = np.zeros(self.nn, dtype=np.int64)
        check(lib().iskra_dsmc_counts_last(self.h, int(did), _p(out)))
        return out

    # surfaces, electrodes, circuit
    def tracker(self, default_surface=None):
        check(lib().iskra_tracker_create(self.h, SURF_ABSORBING if default_surface is None
                                         else int(default_surface)))

    def tracker_surface(self, nodes, surface):
        m = np.ascontiguousarray(np.asarray(nodes, dtype=np.uint8).ravel())
        check(lib().iskra_tracker_surface(self.h, _p(m), int(surface)))

    def tracker_table(self):
        ncell = (self.nodes[0] - 1) * (self.nodes[1] - 1)
        codes = np.zeros(4 * ncell, np.uint8)
        tracked = np.zeros(ncell, np.uint8)
        check(lib().iskra_tracker_table(self.h, _p(codes), _p(tracked)))
        return codes, tracked

    def electrode(self, nodes, fixed=False, sigma=0.0, phi=0.0):
        """nodes: one flag per node, x fastest; returns the electrode id."""
        m = np.ascontiguousarray(np.asarray(nodes, dtype=np.uint8).ravel())
        out = C.c_int()
        check(lib().iskra_electrode_create(self.h, _p(m), int(bool(fixed)), float(sigma),
                                           float(phi), C.byref(out)))
        return out.value

    def electrode_state(self, eid):
        v = [C.c_double() for _ in range(4)]
        check(lib().iskra_electrode_state(self.h, int(eid), *[C.byref(x) for x in v]))
        return dict(sigma=v[0].value, phi=v[1].value, dq=v[2].value, area=v[3].value)

    def circuit(self, R, L, C_, v_offset=0.0, v_amplitude=0.0, v_freq=0.0, v_phase=0.0,
                positive=-1, negative=-1):
        check(lib().iskra_circuit_set(self.h, float(R), float(L), float(C_), float(v_offset),
                                      float(v_amplitude), float(v_freq), float(v_phase),
                                      int(positive), int(negative)))

    def fluid(self, q, m, mu, n):
        """A mobile fluid species (iskra_fluid_add): node densities n, x fastest."""
        v = np.ascontiguousarray(n, dtype=np.float64).ravel()
        assert v.size == self.nn
        out = C.c_int()
        check(lib().iskra_fluid_add(self.h, float(q), float(m), float(mu), _p(v), C.byref(out)))
        return out.value

    def fluid_advance(self, dt):
        check(lib().iskra_fluid_advance(self.h, float(dt)))

    def checkpoint(self):
        """the context's dynamic state as bytes (uint8 array)"""
        n = C.c_int64()
        check(lib().iskra_checkpoint_size(self.h, C.byref(n)))
        buf = np.zeros(max(n.value, 1), dtype=np.uint8)
        check(lib().iskra_checkpoint_save(self.h, _p(buf), n.value))
        return buf[:n.value]

    def restore(self, blob):
        b = np.ascontiguousarray(blob, dtype=np.uint8)
        check(lib().iskra_checkpoint_load(self.h, _p(b), b.size))

    def circuit_state(self):
        v = [C.c_double() for _ in range(4)]
        check(lib().iskra_circuit_state(self.h, *[C.byref(x) for x in v]))
        return dict(i=v[0].value, q=v[1].value, t=v[2].value, vext=v[3].value)

    def rng(self, seed):
        check(lib().iskra_rng_set(self.h, int(seed)))

    def rng_export(self, stream, step, ids, nblocks):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        out = np.zeros(len(ids) * nblocks * 4, dtype=np.uint32)
        check(lib().iskra_rng_export(self.h, stream, step, _p(ids), len(ids), nblocks, _p(out)))
        return out.reshape(len(ids), nblocks, 4)

    def rng_export_noalloc(self, stream, step, ids, nblocks):
        """rng_export staged in the library's fixed per-GPU scratch (no device
        allocation; at most NOALLOC_SCRATCH_BYTES of ids and draws)."""
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        out = np.zeros(len(ids) * nblocks * 4, dtype=np.uint32)
        check(lib().iskra_rng_export_noalloc(self.h, stream, step, _p(ids), len(ids), nblocks,
                                             _p(out)))
        return out.reshape(len(ids), nblocks, 4)

    # loop
    def step(self, dt, nsteps=1):
        check(lib().iskra_step(self.h, float(dt), int(nsteps)))

    def stats(self):
        s = Stats()
        check(lib().iskra_stats_get(self.h, C.byref(s)))
        return dict(steps=s.steps, particle_steps=s.particle_steps,
                    poisson_solves=s.poisson_solves, poisson_iters=s.poisson_iters,
                    last_iters=s.last_iters, last_relres=s.last_relres,
                    kernel_ms={K_NAMES[k]: s.kernel_ms[k] for k in range(7)},
                    kernel_calls={K_NAMES[k]: s.kernel_calls[k] for k in range(7)},
                    push_particles=s.push_particles)

    def profile(self, enable=True):
        check(lib().iskra_profile(self.h, int(enable)))

    def bench_copy(self, nbytes, iters=20, nontemporal=True):
        ms = C.c_double()
        check(lib().iskra_bench_copy(self.h, int(nbytes), int(iters), int(nontemporal),
                                     C.byref(ms)))
        return ms.value

    def sync(self):
        check(lib().iskra_sync(self.h))


class WarmStart:
    """Learned warm start for the 2-D Poisson solve of `ctx` (iskra_ws_*):
    encoder-decoder with `levels` levels and base*2^min(l,3) channels on a
    side x side pixel grid (the context grid has side+1 nodes per axis)."""

    def __init__(self, ctx, levels=4, base=32, side=512, max_batch=8):
        h = P()
        check(lib().iskra_ws_create(ctx.h, levels, base, side, max_batch, C.byref(h)))
        self.h = h
        self.ctx = ctx
        _live_ws.add(self)
        self.levels, self.base, self.side, self.max_batch = levels, base, side, max_batch
        n = C.c_int64()
        check(lib().iskra_ws_param_count(self.h, C.byref(n)))
        self.nparam = n.value

    def close(self):
        if self.h:
            lib().iskra_ws_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        check(lib().iskra_ws_set_params(self.h, _p(p)))

    def get_params(self):
        p = np.zeros(self.nparam, dtype=np.float32)
        check(lib().iskra_ws_get_params(self.h, _p(p)))
        return p

    def set_gain(self, g):
        check(lib().iskra_ws_set_gain(self.h, float(g)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.zeros_like(x)
        check(lib().iskra_ws_forward(self.h, x.shape[0], _p(x), _p(out)))
        return out

    def backward(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = C.c_double()
        check(lib().iskra_ws_backward(self.h, x.shape[0], _p(x), _p(t), _p(g), C.byref(loss)))
        return loss.value, g

    def reserve(self, n):
        check(lib().iskra_ws_dataset_reserve(self.h, int(n)))

    def capture(self):
        i = C.c_int64()
        check(lib().iskra_ws_capture(self.h, C.byref(i)))
        return i.value

    def size(self):
        n = C.c_int64()
        check(lib().iskra_ws_dataset_size(self.h, C.byref(n)))
        return n.value

    def dataset_rms(self, first, count):
        r = C.c_double()
        check(lib().iskra_ws_dataset_rms(self.h, int(first), int(count), C.byref(r)))
        return r.value

    def train(self, first, count, batch, steps, lr, seed=0):
        losses = np.zeros(max(int(steps), 1))
        check(lib().iskra_ws_train(self.h, int(first), int(count), int(batch), int(steps),
                                   float(lr), int(seed), _p(losses)))
        return losses[:steps]

    def eval(self, first, count):
        m = C.c_double()
        check(lib().iskra_ws_eval(self.h, int(first), int(count), C.byref(m)))
        return m.value

    def predict(self):
        check(lib().iskra_ws_predict(self.h))

    def set_objective(self, objective, omega=0.0, sweeps=1):
        """WS_PHI_MSE (warm start) or WS_ENERGY (preconditioner, with the
        Jacobi weight and sweep count the solve will use)."""
        check(lib().iskra_ws_set_objective(self.h, int(objective), float(omega), int(sweeps)))

    def solve(self, warm=WARM_ZERO, omega=0.8, sweeps=1, max_iter=0, raise_unconverged=True):
        """Flexible CG on the context's system preconditioned by the network
        (then `sweeps` damped Jacobi sweeps when omega > 0); returns
        (iters, relres).  raise_unconverged=False returns them also when
        max_iter ran out (ISKRA_E_CONVERGE)."""
        it, rr = C.c_int(), C.c_double()
        rc = lib().iskra_ws_solve(self.h, int(warm), float(omega), int(sweeps), int(max_iter),
                                  C.byref(it), C.byref(rr))
        if rc != E_CONVERGE or raise_unconverged:
            check(rc)
        return it.value, rr.value

    def capture_iterates(self, iters, omega=0.8, sweeps=2):
        """Append the pairs (r, phi - x) of the solve's iterates after each
        count in `iters` (increasing); the context's phi must be the exact
        solution of its rho.  Returns the number appended."""
        a = np.ascontiguousarray(iters, dtype=np.int32)
        n = C.c_int64()
        check(lib().iskra_ws_capture_iterates(self.h, float(omega), int(sweeps), _p(a), a.size,
                                              C.byref(n)))
        return n.value

    def dataset_get(self, index):
        """(input plane, target plane) of dataset sample `index`."""
        a = np.zeros((self.side, self.side), np.float32)
        b = np.zeros((self.side, self.side), np.float32)
        check(lib().iskra_ws_dataset_get(self.h, int(index), _p(a), _p(b)))
        return a, b


class Batch:
    """Contexts advanced together (iskra_batch_*): each member computes
    exactly what it computes when stepped alone."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        arr = (P * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = P()
        check(lib().iskra_batch_create(arr, len(self.ctxs), C.byref(h)))
        self.h = h
        _live_batch.add(self)

    def close(self):
        if self.h:
            lib().iskra_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def graph(self, enable=True):
        check(lib().iskra_batch_set_graph(self.h, int(enable)))

    def step(self, dt, nsteps=1):
        check(lib().iskra_batch_step(self.h, float(dt), int(nsteps)))

    def graphs_launched(self):
        n = C.c_int64()
        check(lib().iskra_batch_graphs_launched(self.h, C.byref(n)))
        return n.value
