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
    "tild_system_load", "tild_system_free", "tild_last_error", "tild_system_info",
    "tild_system_get_positions", "tild_system_get_types", "tild_system_get_molecules",
    "tild_system_get_charges", "tild_system_get_bonds", "tild_system_group_sites",
    "tild_engine_create", "tild_engine_destroy", "tild_engine_last_error",
    "tild_engine_forces", "tild_engine_step", "tild_engine_properties",
    "tild_engine_get_positions", "tild_engine_set_positions", "tild_engine_get_forces",
    "tild_engine_get_density", "tild_engine_get_grid_field", "tild_engine_get_step",
    "tild_engine_synchronize", "tild_engine_get_bins", "tild_engine_profile", "tild_fft_r2c", "tild_fft_c2r",
]


class TildError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Opts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("use_seed", ctypes.c_int), ("seed", ctypes.c_uint32),
                ("spread_mode", ctypes.c_int), ("reserved", ctypes.c_int * 8)]


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
        L.tild_engine_synchronize.argtypes = [vp]
        L.tild_engine_get_bins.argtypes = [vp, ip, sz]
        L.tild_engine_profile.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
        L.tild_fft_r2c.argtypes = [ctypes.c_int, ip, ctypes.c_int, vp, vp, vp]
        L.tild_fft_c2r.argtypes = [ctypes.c_int, ip, ctypes.c_int, vp, vp, vp]
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


class Engine:
    """One device-resident replica (tild_engine)."""

    def __init__(self, system, device=0, seed=None, spread_mode=0, tile_target=0, tile_by=0, tile_bz=0):
        o = Opts()
        o.device = device
        o.spread_mode = spread_mode
        o.reserved[0] = tile_target
        o.reserved[1] = tile_by
        o.reserved[2] = tile_bz
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

    def grid_field(self, t, d):
        a = np.zeros(self.M, dtype=np.float32)
        self._chk(lib().tild_engine_get_grid_field(self._h, t, d, _fp(a), a.size), "get_grid_field")
        return a.reshape(self.grid_shape)

    def bins(self):
        a = np.zeros((self.ns, self.dim), dtype=np.int32)
        self._chk(lib().tild_engine_get_bins(self._h, _ip(a), a.size), "get_bins")
        return a

    PHASES = ["spread", "x_r2c", "y_fft", "last_fused", "y_ifft", "x_c2r", "particles", "advance"]

    def profile(self, nsteps):
        a = np.zeros(8, dtype=np.float64)
        self._chk(lib().tild_engine_profile(self._h, int(nsteps), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 8),
                  "profile")
        return dict(zip(self.PHASES, a.tolist()))

    def step_count(self):
        s = ctypes.c_int()
        self._chk(lib().tild_engine_get_step(self._h, ctypes.byref(s)), "get_step")
        return s.value

    def synchronize(self):
        self._chk(lib().tild_engine_synchronize(self._h), "synchronize")


def fft_r2c(dim, Nx, batch, d_in, d_out, stream=None):
    arr = (ctypes.c_int * 3)(*(list(Nx) + [1] * (3 - len(Nx))))
    _check(lib().tild_fft_r2c(dim, arr, batch, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                              ctypes.c_void_p(stream) if stream else None), "tild_fft_r2c")


def fft_c2r(dim, Nx, batch, d_in, d_out, stream=None):
    arr = (ctypes.c_int * 3)(*(list(Nx) + [1] * (3 - len(Nx))))
    _check(lib().tild_fft_c2r(dim, arr, batch, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                              ctypes.c_void_p(stream) if stream else None), "tild_fft_c2r")
