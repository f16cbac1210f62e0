"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

Test infrastructure: only tests/, the smoke check and bench.py's CPU-baseline leg use
this module, always as the checker.
"""
import ctypes
import os
import pathlib

import numpy as np

REPO = pathlib.Path(__file__).resolve().parent.parent
LIB_PATH = REPO / "oracle" / "liboracle.so"

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run build())")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        u32 = ctypes.c_uint32
        L.ora_create.restype = vp
        L.ora_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, u32, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.ora_destroy.argtypes = [vp]
        L.ora_get_int.restype = ctypes.c_int
        L.ora_get_int.argtypes = [vp, ctypes.c_char_p]
        L.ora_get_double.restype = ctypes.c_double
        L.ora_get_double.argtypes = [vp, ctypes.c_char_p]
        for name in ("ora_get_positions", "ora_set_positions", "ora_get_forces", "ora_get_charges",
                     "ora_get_charge_density"):
            getattr(L, name).argtypes = [vp, dp]
        for name in ("ora_get_types", "ora_get_molecules"):
            getattr(L, name).argtypes = [vp, ip]
        L.ora_get_density.argtypes = [vp, ctypes.c_int, dp]
        L.ora_get_grid_force.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.ora_get_electric_field.argtypes = [vp, ctypes.c_int, dp]
        L.ora_get_grid_stencil.argtypes = [vp, ip, dp]
        L.ora_get_bonds.argtypes = [vp, ip, ip, ip]
        L.ora_group_sites.restype = ctypes.c_int
        L.ora_group_sites.argtypes = [vp, ctypes.c_int, ip]
        L.ora_forces.argtypes = [vp]
        L.ora_integrate.argtypes = [vp]
        L.ora_run.argtypes = [vp, ctypes.c_int]
        L.ora_properties.restype = ctypes.c_int
        L.ora_properties.argtypes = [vp, dp, ctypes.c_int]
        L.ora_philox4x32_10.argtypes = [ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)]
        L.ora_normals.argtypes = [u32, u32, u32, u32, dp]
        L.ora_fft.restype = ctypes.c_int
        L.ora_fft.argtypes = [dp, ctypes.c_int, ip, ctypes.c_int]
        L.ora_fft1d.restype = ctypes.c_int
        L.ora_fft1d.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.ora_set_double.restype = ctypes.c_int
        L.ora_set_double.argtypes = [vp, ctypes.c_char_p, ctypes.c_double]
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_get_threads.restype = ctypes.c_int
        L.ora_sk_accumulate.argtypes = [vp, ctypes.c_int]
        L.ora_sk_average.restype = ctypes.c_int
        L.ora_sk_average.argtypes = [vp, ctypes.c_int, dp]
        L.ora_nlist.restype = ctypes.c_int
        L.ora_nlist.argtypes = [vp, ctypes.c_double, ctypes.c_int, ip, ip, ip, ctypes.c_int, ip, ip]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def philox(ctr, key):
    L = lib()
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    L.ora_philox4x32_10(c, k, o)
    return list(o)


def normals(seed, particle, step, stream=0):
    z = np.zeros(4)
    lib().ora_normals(seed, particle, step, stream, _dp(z))
    return z


def set_threads(n):
    """Worker threads of the oracle's parallel loops (results do not depend on it)."""
    lib().ora_set_threads(int(n))


def get_threads():
    return lib().ora_get_threads()


def fft(a, sign=-1):
    """Unnormalised C2C FFT of a complex grid given in (z,y,x) numpy order."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    dim = a.ndim
    nx = (ctypes.c_int * 3)(*(list(a.shape[::-1]) + [1] * (3 - dim)))
    buf = a.copy()
    rc = lib().ora_fft(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), dim, nx, sign)
    if rc:
        raise RuntimeError(f"ora_fft rc={rc}")
    return buf


class System:
    """One oracle system built from an input deck (text) and a base directory."""

    def __init__(self, input_text, base_dir, seed=None):
        L = lib()
        err = ctypes.create_string_buffer(512)
        self._h = L.ora_create(input_text.encode(), str(base_dir).encode(),
                               0 if seed is None else int(seed) & 0xFFFFFFFF,
                               0 if seed is None else 1, err, 512)
        if not self._h:
            raise ValueError(err.value.decode())
        self.ns = L.ora_get_int(self._h, b"ns")
        self.dim = L.ora_get_int(self._h, b"dim")
        self.ntypes = L.ora_get_int(self._h, b"ntypes")
        self.M = L.ora_get_int(self._h, b"M")
        self.S = L.ora_get_int(self._h, b"S")
        self.Nx = [L.ora_get_int(self._h, f"Nx{i}".encode()) for i in range(3)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ora_destroy(h)
            self._h = None

    def geti(self, k):
        return lib().ora_get_int(self._h, k.encode())

    def getd(self, k):
        return lib().ora_get_double(self._h, k.encode())

    def set_noise(self, mag):
        """Test-only: thermal-noise amplitude of the EM/GJF updates (0 = off)."""
        if lib().ora_set_double(self._h, b"noise_mag", float(mag)):
            raise ValueError("noise_mag not settable")

    @property
    def grid_shape(self):
        return tuple(self.Nx[:self.dim][::-1])

    def positions(self):
        a = np.zeros((self.ns, self.dim))
        lib().ora_get_positions(self._h, _dp(a))
        return a

    def set_positions(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64).reshape(self.ns, self.dim)
        lib().ora_set_positions(self._h, _dp(a))

    def types(self):
        a = np.zeros(self.ns, dtype=np.int32)
        lib().ora_get_types(self._h, _ip(a))
        return a

    def molecules(self):
        a = np.zeros(self.ns, dtype=np.int32)
        lib().ora_get_molecules(self._h, _ip(a))
        return a

    def charges(self):
        a = np.zeros(self.ns)
        lib().ora_get_charges(self._h, _dp(a))
        return a

    def bonds(self):
        nb = np.zeros(self.ns, dtype=np.int32)
        bt = np.zeros((self.ns, 3), dtype=np.int32)
        ty = np.zeros((self.ns, 3), dtype=np.int32)
        lib().ora_get_bonds(self._h, _ip(nb), _ip(bt), _ip(ty))
        return nb, bt, ty

    def group_sites(self, g):
        n = lib().ora_group_sites(self._h, g, None)
        a = np.zeros(max(n, 1), dtype=np.int32)
        lib().ora_group_sites(self._h, g, _ip(a))
        return a[:n]

    def forces(self):
        lib().ora_forces(self._h)
        a = np.zeros((self.ns, self.dim))
        lib().ora_get_forces(self._h, _dp(a))
        return a

    def last_forces(self):
        a = np.zeros((self.ns, self.dim))
        lib().ora_get_forces(self._h, _dp(a))
        return a

    def density(self, t):
        a = np.zeros(self.M)
        lib().ora_get_density(self._h, t, _dp(a))
        return a.reshape(self.grid_shape)

    def charge_density(self):
        a = np.zeros(self.M)
        lib().ora_get_charge_density(self._h, _dp(a))
        return a.reshape(self.grid_shape)

    def grid_force(self, t, d):
        a = np.zeros(self.M)
        lib().ora_get_grid_force(self._h, t, d, _dp(a))
        return a.reshape(self.grid_shape)

    def electric_field(self, d):
        a = np.zeros(self.M)
        lib().ora_get_electric_field(self._h, d, _dp(a))
        return a.reshape(self.grid_shape)

    def stencil(self):
        inds = np.zeros((self.ns, self.S), dtype=np.int32)
        W = np.zeros((self.ns, self.S))
        lib().ora_get_grid_stencil(self._h, _ip(inds), _dp(W))
        return inds, W

    def properties(self):
        out = np.zeros(9 + 256)
        n = lib().ora_properties(self._h, _dp(out), len(out))
        v = out[:n]
        return {"Upe": v[0], "Ubond": v[1], "Uelec": v[2], "Ptens": v[3:9].copy(),
                "U_styles": v[9:].copy()}

    def sk_accumulate(self, t):
        lib().ora_sk_accumulate(self._h, t)

    def sk_average(self, t):
        a = np.zeros(2 * self.M)
        n = lib().ora_sk_average(self._h, t, _dp(a))
        return (a[0::2] + 1j * a[1::2]).reshape(self.grid_shape), n

    def integrate(self):
        lib().ora_integrate(self._h)

    def run(self, n):
        lib().ora_run(self._h, n)

    def nlist(self, rc, group=0):
        cap = 4096
        cell_of = np.zeros(self.ns, dtype=np.int32)
        counts = np.zeros(self.ns, dtype=np.int32)
        neigh = np.zeros((self.ns, cap), dtype=np.int32)
        cmax = ctypes.c_int(0)
        nc = ctypes.c_int(0)
        rc_ = lib().ora_nlist(self._h, rc, group, _ip(cell_of), _ip(counts), _ip(neigh), cap,
                              ctypes.byref(cmax), ctypes.byref(nc))
        if rc_:
            raise RuntimeError(f"ora_nlist rc={rc_}")
        return cell_of, counts, neigh, cmax.value, nc.value
