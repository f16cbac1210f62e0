"""ctypes binding of libtild.so (the C ABI declared in include/tild.h).

The binding loads the in-tree library built by the package Makefile and raises if it
is missing: there is no fallback implementation.
"""
import ctypes
import pathlib

import numpy as np

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libtild.so"

TILD_MAX_STYLES = 64
STATUS = {0: "OK", 1: "PARSE", 2: "CONFIG", 3: "OOM", 4: "DEVICE", 5: "NAN", 6: "ARG",
          7: "BUDGET", 8: "UNSUPPORTED"}

# Every entry point include/tild.h declares (checked by the CPU tests).
EXPORTS = [
    "tild_system_load", "tild_system_free", "tild_last_error", "tild_system_info", "tild_system_info_text",
    "tild_system_get_positions", "tild_system_get_types", "tild_system_get_molecules",
    "tild_system_get_charges", "tild_system_get_bonds", "tild_system_group_sites",
    "tild_engine_create", "tild_engine_destroy", "tild_engine_last_error",
    "tild_engine_forces", "tild_engine_step", "tild_engine_properties",
    "tild_engine_get_positions", "tild_engine_set_positions", "tild_engine_get_forces",
    "tild_engine_get_density", "tild_engine_get_grid_field", "tild_engine_get_step",
    "tild_engine_get_state", "tild_engine_set_state",
    "tild_engine_synchronize", "tild_engine_get_bins", "tild_engine_profile", "tild_engine_nlist", "tild_fft_r2c", "tild_fft_c2r",
    "tild_fft_c2r_noalloc", "tild_system_compute", "tild_engine_sk_accumulate", "tild_engine_sk_average",
    "tild_engine_geometry", "tild_batch_geometry",
    "tild_batch_create", "tild_batch_destroy", "tild_batch_last_error", "tild_batch_add", "tild_batch_remove",
    "tild_batch_step", "tild_batch_properties", "tild_batch_get_positions", "tild_batch_set_positions",
    "tild_batch_get_density", "tild_batch_replica_step", "tild_batch_profile",
    "tild_server_create", "tild_server_destroy", "tild_server_submit", "tild_server_poll", "tild_server_wait",
    "tild_server_drain", "tild_server_result", "tild_server_stats_get",
]

JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class TildError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("use_seed", ctypes.c_int), ("seed", ctypes.c_uint32),
                ("spread_mode", ctypes.c_int), ("tile_target", ctypes.c_int), ("tile_by", ctypes.c_int),
                ("tile_bz", ctypes.c_int), ("reorder_every", ctypes.c_int), ("fft_mid_elems", ctypes.c_int),
                ("fft_last_lines", ctypes.c_int), ("fft_c2r_rows", ctypes.c_int), ("fft_plane", ctypes.c_int),
                ("fft_last_dirs", ctypes.c_int), ("tile_slack", ctypes.c_int), ("autotune", ctypes.c_int),
                ("deposit_threads", ctypes.c_int), ("fft_staged", ctypes.c_int)]


class Geometry(ctypes.Structure):
    _fields_ = [("tile_target", ctypes.c_int), ("tiles", ctypes.c_int), ("deposit_threads", ctypes.c_int),
                ("plane", ctypes.c_int), ("last_lines", ctypes.c_int), ("last_dirs", ctypes.c_int),
                ("mid_elems", ctypes.c_int), ("c2r_rows", ctypes.c_int), ("last_reg", ctypes.c_int),
                ("grad_reg", ctypes.c_int), ("plane_reg", ctypes.c_int), ("plane_inv", ctypes.c_int),
                ("tuned", ctypes.c_int),
                ("from_cache", ctypes.c_int), ("step_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Props(ctypes.Structure):
    _fields_ = [("Upe", ctypes.c_double), ("Ubond", ctypes.c_double), ("Uelec", ctypes.c_double),
                ("Ptens", ctypes.c_double * 6), ("n_styles", ctypes.c_int),
                ("U_style", ctypes.c_double * TILD_MAX_STYLES), ("step", ctypes.c_int),
                ("nan_flag", ctypes.c_int)]

    def as_dict(self):
        return {"Upe": self.Upe, "Ubond": self.Ubond, "Uelec": self.Uelec,
                "Ptens": np.array(self.Ptens[:]), "U_styles": np.array(self.U_style[:self.n_styles]),
                "step": self.step, "nan_flag": self.nan_flag}


_lib = None


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("chunk_steps", ctypes.c_int), ("max_replicas", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("hbm_budget", ctypes.c_size_t), ("step_work", ctypes.c_int),
                ("reserved", ctypes.c_int * 7)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int), ("queued", ctypes.c_int), ("running", ctypes.c_int),
                ("done", ctypes.c_int), ("failed", ctypes.c_int), ("bytes_in_use", ctypes.c_size_t),
                ("peak_bytes", ctypes.c_size_t), ("budget", ctypes.c_size_t), ("busy_s", ctypes.c_double),
                ("elapsed_s", ctypes.c_double), ("bead_steps", ctypes.c_double), ("p50_s", ctypes.c_double),
                ("p99_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("retire_s", ctypes.c_double),
                ("wait50_s", ctypes.c_double), ("wait99_s", ctypes.c_double), ("svc50_s", ctypes.c_double),
                ("svc99_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libtild.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        fp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
        L.tild_last_error.restype = ctypes.c_char_p
        L.tild_system_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(vp)]
        L.tild_system_free.argtypes = [vp]
        L.tild_system_info.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
        L.tild_system_info_text.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, sz]
        L.tild_system_get_positions.argtypes = [vp, fp, sz]
        L.tild_system_get_types.argtypes = [vp, ip, sz]
        L.tild_system_get_molecules.argtypes = [vp, ip, sz]
        L.tild_system_get_charges.argtypes = [vp, fp, sz]
        L.tild_system_get_bonds.argtypes = [vp, ip, ip, ip, sz]
        L.tild_system_group_sites.argtypes = [vp, ctypes.c_int, ip, sz, ip]
        L.tild_engine_create.argtypes = [vp, ctypes.POINTER(Opts), ctypes.POINTER(vp)]
        L.tild_engine_destroy.argtypes = [vp]
        L.tild_engine_last_error.restype = ctypes.c_char_p
        L.tild_engine_last_error.argtypes = [vp]
        L.tild_engine_forces.argtypes = [vp]
        L.tild_engine_step.argtypes = [vp, ctypes.c_int]
        L.tild_engine_properties.argtypes = [vp, ctypes.POINTER(Props)]
        L.tild_engine_get_positions.argtypes = [vp, fp, sz]
        L.tild_engine_set_positions.argtypes = [vp, fp, sz]
        L.tild_engine_get_forces.argtypes = [vp, fp, sz]
        L.tild_engine_get_density.argtypes = [vp, ctypes.c_int, fp, sz]
        L.tild_engine_get_grid_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, fp, sz]
        L.tild_engine_get_step.argtypes = [vp, ip]
        L.tild_engine_get_state.argtypes = [vp, ip, fp, sz]
        L.tild_engine_set_state.argtypes = [vp, ctypes.c_int, fp, sz]
        L.tild_engine_synchronize.argtypes = [vp]
        L.tild_engine_get_bins.argtypes = [vp, ip, sz]
        L.tild_engine_profile.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.tild_engine_nlist.argtypes = [vp, ctypes.c_float, ip, ctypes.c_int, ip, ip, ip, sz, ip, ip]
        L.tild_batch_create.argtypes = [ctypes.POINTER(Opts), ctypes.POINTER(vp)]
        L.tild_batch_destroy.argtypes = [vp]
        L.tild_batch_last_error.restype = ctypes.c_char_p
        L.tild_batch_last_error.argtypes = [vp]
        L.tild_batch_add.argtypes = [vp, vp, ctypes.c_uint32, ip]
        L.tild_batch_remove.argtypes = [vp, ctypes.c_int]
        L.tild_batch_step.argtypes = [vp, ip, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.tild_batch_properties.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Props)]
        L.tild_batch_get_positions.argtypes = [vp, ctypes.c_int, fp, sz]
        L.tild_batch_set_positions.argtypes = [vp, ctypes.c_int, fp, sz]
        L.tild_batch_get_density.argtypes = [vp, ctypes.c_int, ctypes.c_int, fp, sz]
        L.tild_batch_replica_step.argtypes = [vp, ctypes.c_int, ip]
        L.tild_batch_profile.argtypes = [vp, ip, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.tild_server_create.argtypes = [ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)]
        L.tild_server_destroy.argtypes = [vp]
        L.tild_server_submit.argtypes = [vp, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_int, ip]
        L.tild_server_poll.argtypes = [vp, ctypes.c_int, ip]
        L.tild_server_wait.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.tild_server_drain.argtypes = [vp, ctypes.c_double]
        L.tild_server_result.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Props), fp, sz,
                                         ctypes.POINTER(ctypes.c_double)]
        L.tild_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.tild_fft_r2c.argtypes = [ctypes.c_int, ip, ctypes.c_int, vp, vp, vp]
        L.tild_fft_c2r.argtypes = [ctypes.c_int, ip, ctypes.c_int, vp, vp, vp]
        L.tild_fft_c2r_noalloc.argtypes = [ctypes.c_int, ip, ctypes.c_int, vp, vp, vp]
        L.tild_system_compute.argtypes = [vp, ctypes.c_int, ip, ip, ip]
        L.tild_engine_sk_accumulate.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.tild_engine_sk_average.argtypes = [vp, ctypes.c_int, fp, sz, ip]
        L.tild_engine_geometry.argtypes = [vp, vp]
        L.tild_batch_geometry.argtypes = [vp, ctypes.c_int, vp]
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _check(st, where=""):
    if st != 0:
        raise TildError(st, f"{where}: {lib().tild_last_error().decode()}")


class System:
    """A parsed deck (host only; no GPU needed)."""

    def __init__(self, input_text, base_dir="."):
        h = ctypes.c_void_p()
        _check(lib().tild_system_load(input_text.encode(), str(base_dir).encode(), ctypes.byref(h)),
               "tild_system_load")
        self._h = h
        self.ns = int(self.info("ns"))
        self.dim = int(self.info("dim"))
        self.ntypes = int(self.info("ntypes"))
        self.M = int(self.info("M"))
        self.Nx = [int(self.info(f"Nx{i}")) for i in range(3)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().tild_system_free(h)
            self._h = None

    @property
    def grid_shape(self):
        return tuple(self.Nx[:self.dim][::-1])

    def info(self, key):
        v = ctypes.c_double()
        _check(lib().tild_system_info(self._h, key.encode(), ctypes.byref(v)), "tild_system_info")
        return v.value

    def positions(self):
        a = np.zeros((self.ns, self.dim), dtype=np.float32)
        _check(lib().tild_system_get_positions(self._h, _fp(a), a.size))
        return a

    def types(self):
        a = np.zeros(self.ns, dtype=np.int32)
        _check(lib().tild_system_get_types(self._h, _ip(a), a.size))
        return a

    def molecules(self):
        a = np.zeros(self.ns, dtype=np.int32)
        _check(lib().tild_system_get_molecules(self._h, _ip(a), a.size))
        return a

    def charges(self):
        a = np.zeros(self.ns, dtype=np.float32)
        _check(lib().tild_system_get_charges(self._h, _fp(a), a.size))
        return a

    def bonds(self):
        nb = np.zeros(self.ns, dtype=np.int32)
        bt = np.zeros((self.ns, 3), dtype=np.int32)
        ty = np.zeros((self.ns, 3), dtype=np.int32)
        _check(lib().tild_system_get_bonds(self._h, _ip(nb), _ip(bt), _ip(ty), self.ns))
        return nb, bt, ty

    def group_sites(self, g):
        n = ctypes.c_int()
        _check(lib().tild_system_group_sites(self._h, g, None, 0, ctypes.byref(n)))
        a = np.zeros(max(n.value, 1), dtype=np.int32)
        _check(lib().tild_system_group_sites(self._h, g, _ip(a), a.size, ctypes.byref(n)))
        return a[:n.value]

    def computes(self):
        """The deck's avg_sk computes as (0-based type, freq, wait) tuples."""
        out = []
        for i in range(int(self.info("n_computes"))):
            t, f, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _check(lib().tild_system_compute(self._h, i, ctypes.byref(t), ctypes.byref(f), ctypes.byref(w)))
            out.append((t.value, f.value, w.value))
        return out


class Engine:
    """One device-resident replica (tild_engine)."""

    def __init__(self, system, device=0, seed=None, spread_mode=0, tile_target=0, tile_by=0, tile_bz=0,
                 reorder_every=0, fft_mid_elems=0, fft_last_lines=0, fft_c2r_rows=0, fft_plane=0,
                 fft_last_dirs=0, tile_slack=0, autotune=False, deposit_threads=0, fft_staged=False):
        o = Opts()
        o.device = device
        o.spread_mode = spread_mode
        o.autotune, o.deposit_threads = int(bool(autotune)), deposit_threads
        o.tile_target, o.tile_by, o.tile_bz = tile_target, tile_by, tile_bz
        o.reorder_every = reorder_every
        o.fft_mid_elems, o.fft_last_lines, o.fft_c2r_rows = fft_mid_elems, fft_last_lines, fft_c2r_rows
        o.fft_plane, o.fft_last_dirs = fft_plane, fft_last_dirs
        o.fft_staged = int(bool(fft_staged))
        o.tile_slack = tile_slack
        if seed is not None:
            o.use_seed = 1
            o.seed = int(seed) & 0xFFFFFFFF
        h = ctypes.c_void_p()
        _check(lib().tild_engine_create(system._h, ctypes.byref(o), ctypes.byref(h)), "tild_engine_create")
        self._h = h
        self.ns, self.dim, self.ntypes, self.M = system.ns, system.dim, system.ntypes, system.M
        self.grid_shape = system.grid_shape
        self.nfields = self.ntypes + (1 if system.info("do_charges") else 0)

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().tild_engine_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, st, where):
        if st != 0:
            raise TildError(st, f"{where}: {lib().tild_engine_last_error(self._h).decode()}")

    def forces(self):
        self._chk(lib().tild_engine_forces(self._h), "forces")
        return self.get_forces()

    def step(self, n=1):
        self._chk(lib().tild_engine_step(self._h, int(n)), "step")

    def properties(self):
        p = Props()
        self._chk(lib().tild_engine_properties(self._h, ctypes.byref(p)), "properties")
        return p.as_dict()

    def positions(self):
        a = np.zeros((self.ns, self.dim), dtype=np.float32)
        self._chk(lib().tild_engine_get_positions(self._h, _fp(a), a.size), "get_positions")
        return a

    def set_positions(self, x):
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(self.ns, self.dim)
        self._chk(lib().tild_engine_set_positions(self._h, _fp(a), a.size), "set_positions")

    def get_forces(self):
        a = np.zeros((self.ns, self.dim), dtype=np.float32)
        self._chk(lib().tild_engine_get_forces(self._h, _fp(a), a.size), "get_forces")
        return a

    def density(self, t):
        a = np.zeros(self.M, dtype=np.float32)
        self._chk(lib().tild_engine_get_density(self._h, t, _fp(a), a.size), "get_density")
        return a.reshape(self.grid_shape)

    def geometry(self):
        """Block geometry in use (and the autotuner's pick, if tuned)."""
        g = Geometry()
        self._chk(lib().tild_engine_geometry(self._h, ctypes.byref(g)), "geometry")
        return g.as_dict()

    def sk_accumulate(self, t, slot=None):
        """avg_sk: add rho_t(k) rho_t(k) / M of the last force evaluation's densities to
        compute slot `slot` (default: slot t)."""
        self._chk(lib().tild_engine_sk_accumulate(self._h, t if slot is None else slot, t), "sk_accumulate")

    def sk_average(self, slot):
        """avg_sk running average of a compute slot over all M wavevectors (complex,
        grid shape) and the number of evaluations."""
        a = np.zeros(2 * self.M, dtype=np.float32)
        n = ctypes.c_int()
        self._chk(lib().tild_engine_sk_average(self._h, slot, _fp(a), a.size, ctypes.byref(n)), "sk_average")
        return (a[0::2] + 1j * a[1::2]).astype(np.complex64).reshape(self.grid_shape), n.value

    def grid_field(self, t, d):
        a = np.zeros(self.M, dtype=np.float32)
        self._chk(lib().tild_engine_get_grid_field(self._h, t, d, _fp(a), a.size), "get_grid_field")
        return a.reshape(self.grid_shape)

    def bins(self):
        a = np.zeros((self.ns, self.dim), dtype=np.int32)
        self._chk(lib().tild_engine_get_bins(self._h, _ip(a), a.size), "get_bins")
        return a

    PHASES = ["claim", "deposit", "plane_r2c", "last_fused", "y_ifft", "x_c2r", "particles", "advance", "reorder"]

    def profile(self, nsteps):
        n = len(self.PHASES)
        a = np.zeros(n, dtype=np.float64)
        self._chk(lib().tild_engine_profile(self._h, int(nsteps), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n),
                  "profile")
        return dict(zip(self.PHASES, a.tolist()))

    def nlist(self, rc, sites=None, cap=None):
        """(cell_of, counts, neigh (ns, cap), cell_max, ncells) by site id; cap defaults to
        the reference's CELL_MAX."""
        cm, nc = ctypes.c_int(), ctypes.c_int()
        sarr = None if sites is None else np.ascontiguousarray(sites, dtype=np.int32)
        sp = _ip(sarr) if sarr is not None else None
        nsites = 0 if sarr is None else len(sarr)
        self._chk(lib().tild_engine_nlist(self._h, float(rc), sp, nsites, None, None, None, 0,
                                          ctypes.byref(cm), ctypes.byref(nc)), "nlist size")
        cap = max(cm.value, 1) if cap is None else int(cap)
        cell_of = np.zeros(self.ns, dtype=np.int32)
        counts = np.zeros(self.ns, dtype=np.int32)
        neigh = np.full((self.ns, cap), -1, dtype=np.int32)
        self._chk(lib().tild_engine_nlist(self._h, float(rc), sp, nsites, _ip(cell_of), _ip(counts), _ip(neigh), cap,
                                          ctypes.byref(cm), ctypes.byref(nc)), "nlist")
        return cell_of, counts, neigh, cm.value, nc.value

    def step_count(self):
        s = ctypes.c_int()
        self._chk(lib().tild_engine_get_step(self._h, ctypes.byref(s)), "get_step")
        return s.value

    def state(self):
        """(step, GJF displacement x - xo per site): with positions() and the seed, an
        exact restart point."""
        s = ctypes.c_int()
        a = np.zeros((self.ns, self.dim), dtype=np.float32)
        self._chk(lib().tild_engine_get_state(self._h, ctypes.byref(s), _fp(a), a.size), "get_state")
        return s.value, a

    def set_state(self, step, xo_disp):
        a = np.ascontiguousarray(xo_disp, dtype=np.float32).reshape(self.ns, self.dim)
        self._chk(lib().tild_engine_set_state(self._h, int(step), _fp(a), a.size), "set_state")

    def synchronize(self):
        self._chk(lib().tild_engine_synchronize(self._h), "synchronize")


class Batch:
    """Many device-resident replicas advanced together (tild_batch)."""

    def __init__(self, device=0, spread_mode=0, tile_target=0, reorder_every=0, autotune=False, deposit_threads=0,
                 fft_staged=False):
        o = Opts()
        o.device = device
        o.spread_mode = spread_mode
        o.autotune, o.deposit_threads = int(bool(autotune)), deposit_threads
        o.fft_staged = int(bool(fft_staged))
        o.tile_target = tile_target
        o.reorder_every = reorder_every
        h = ctypes.c_void_p()
        _check(lib().tild_batch_create(ctypes.byref(o), ctypes.byref(h)), "tild_batch_create")
        self._h = h
        self.meta = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().tild_batch_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, st, where):
        if st != 0:
            raise TildError(st, f"{where}: {lib().tild_batch_last_error(self._h).decode()}")

    def geometry(self, r):
        """Block geometry of replica r (and the autotuner's pick, if tuned)."""
        g = Geometry()
        self._chk(lib().tild_batch_geometry(self._h, r, ctypes.byref(g)), "geometry")
        return g.as_dict()

    def add(self, system, seed):
        r = ctypes.c_int()
        self._chk(lib().tild_batch_add(self._h, system._h, int(seed) & 0xFFFFFFFF, ctypes.byref(r)), "add")
        self.meta[r.value] = (system.ns, system.dim, system.M, system.grid_shape)
        return r.value

    def remove(self, r):
        self._chk(lib().tild_batch_remove(self._h, r), "remove")
        self.meta.pop(r, None)

    def step(self, n, replicas=None, graph=True):
        ids = list(replicas) if replicas is not None else []
        arr = (ctypes.c_int * max(1, len(ids)))(*ids)
        self._chk(lib().tild_batch_step(self._h, arr if ids else None, len(ids), int(n), 1 if graph else 0), "step")

    def properties(self, r):
        p = Props()
        self._chk(lib().tild_batch_properties(self._h, r, ctypes.byref(p)), "properties")
        return p.as_dict()

    def positions(self, r):
        ns, dim = self.meta[r][:2]
        a = np.zeros((ns, dim), dtype=np.float32)
        self._chk(lib().tild_batch_get_positions(self._h, r, _fp(a), a.size), "get_positions")
        return a

    def set_positions(self, r, x):
        ns, dim = self.meta[r][:2]
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(ns, dim)
        self._chk(lib().tild_batch_set_positions(self._h, r, _fp(a), a.size), "set_positions")

    def density(self, r, t):
        M, shape = self.meta[r][2], self.meta[r][3]
        a = np.zeros(M, dtype=np.float32)
        self._chk(lib().tild_batch_get_density(self._h, r, t, _fp(a), a.size), "get_density")
        return a.reshape(shape)

    def replica_step(self, r):
        s = ctypes.c_int()
        self._chk(lib().tild_batch_replica_step(self._h, r, ctypes.byref(s)), "replica_step")
        return s.value

    def profile(self, nsteps, replicas=None):
        ids = list(replicas) if replicas is not None else list(self.meta)
        arr = (ctypes.c_int * max(1, len(ids)))(*ids)
        n = len(Engine.PHASES)
        a = np.zeros(n, dtype=np.float64)
        self._chk(lib().tild_batch_profile(self._h, arr, len(ids), int(nsteps),
                                           a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n), "profile")
        return dict(zip(Engine.PHASES, a.tolist()))


class Server:
    """Continuous-batching job server on one device (tild_server).

    submit() parses the deck on the calling thread and queues the job; a worker
    thread inside the library admits, steps and retires jobs.  The ctypes calls
    release the GIL, so wait()/drain() block only the calling thread."""

    def __init__(self, device=0, chunk_steps=100, max_replicas=0, use_graph=True, hbm_budget=0, step_work=0):
        o = ServerOpts()
        o.device, o.chunk_steps, o.max_replicas = device, chunk_steps, max_replicas
        o.use_graph, o.hbm_budget, o.step_work = 1 if use_graph else 0, int(hbm_budget), int(step_work)
        h = ctypes.c_void_p()
        _check(lib().tild_server_create(ctypes.byref(o), ctypes.byref(h)), "tild_server_create")
        self._h = h
        self.meta = {}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().tild_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, input_text, base_dir, seed, nsteps, ns=None, dim=None):
        j = ctypes.c_int()
        _check(lib().tild_server_submit(self._h, input_text.encode(), str(base_dir).encode(),
                                        int(seed) & 0xFFFFFFFF, int(nsteps), ctypes.byref(j)), "submit")
        self.meta[j.value] = (ns, dim)
        return j.value

    def poll(self, job):
        s = ctypes.c_int()
        _check(lib().tild_server_poll(self._h, job, ctypes.byref(s)), "poll")
        return JOB_STATES[s.value]

    def wait(self, job, timeout=-1.0):
        _check(lib().tild_server_wait(self._h, job, float(timeout)), "wait")

    def drain(self, timeout=-1.0):
        _check(lib().tild_server_drain(self._h, float(timeout)), "drain")

    def result(self, job, positions=True):
        """(properties dict, positions (ns, dim) or None, seconds from submission to result)."""
        p = Props()
        lat = ctypes.c_double()
        ns, dim = self.meta.get(job, (None, None))
        x = None
        if positions and ns:
            x = np.zeros((ns, dim), dtype=np.float32)
        _check(lib().tild_server_result(self._h, job, ctypes.byref(p), _fp(x) if x is not None else None,
                                        x.size if x is not None else 0, ctypes.byref(lat)), "result")
        return p.as_dict(), x, lat.value

    def stats(self):
        st = ServerStats()
        _check(lib().tild_server_stats_get(self._h, ctypes.byref(st)), "stats")
        return st.as_dict()


def fft_r2c(dim, Nx, batch, d_in, d_out, stream=None):
    arr = (ctypes.c_int * 3)(*(list(Nx) + [1] * (3 - len(Nx))))
    _check(lib().tild_fft_r2c(dim, arr, batch, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                              ctypes.c_void_p(stream) if stream else None), "tild_fft_r2c")


def fft_c2r(dim, Nx, batch, d_in, d_out, stream=None):
    arr = (ctypes.c_int * 3)(*(list(Nx) + [1] * (3 - len(Nx))))
    _check(lib().tild_fft_c2r(dim, arr, batch, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                              ctypes.c_void_p(stream) if stream else None), "tild_fft_c2r")


def fft_c2r_noalloc(dim, Nx, batch, d_in, d_out, stream=None):
    """fft_c2r on the library's fixed device scratch: no device allocation."""
    arr = (ctypes.c_int * 3)(*(list(Nx) + [1] * (3 - len(Nx))))
    _check(lib().tild_fft_c2r_noalloc(dim, arr, batch, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                                      ctypes.c_void_p(stream) if stream else None), "tild_fft_c2r_noalloc")
