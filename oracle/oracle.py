"""ctypes wrapper of the CPU oracle (liboracle.so, built from fdtd_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker or the timed CPU
baseline.  The product package fdtd_amd never imports it.

Field ids follow the column order of the reference's probe writer
(fdtd.cu:1223): 0..2 Dx Dy Dz, 3..5 Hx Hy Hz, 6..8 Ex Ey Ez.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

DEBYE_REF = 0
DEBYE_FIXED = 1
FIELDS = ["Dx", "Dy", "Dz", "Hx", "Hy", "Hz", "Ex", "Ey", "Ez"]

_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("iterations", C.c_int), ("epw", C.c_int),
        ("wave_frequency", C.c_float), ("pulse_width", C.c_float),
        ("pulse_modulation_frequency", C.c_float),
        ("nsrc", C.c_int), ("src", C.c_int * 192),
        ("sigma", C.c_float), ("eps_s", C.c_float), ("eps_i", C.c_float), ("tau_d", C.c_float),
        ("pi", C.c_float), ("c", C.c_float), ("timeskip", C.c_float), ("lambda_", C.c_float),
        ("dt", C.c_float), ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("mu0", C.c_float), ("eps0", C.c_float),
        ("input_path", C.c_char * 1024), ("output_path", C.c_char * 1024),
    ]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dt", C.c_float), ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("mu0", C.c_float), ("eps0", C.c_float)]


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_parse_params.argtypes = [C.c_char_p, C.POINTER(Params)]
        L.or_derive_params.argtypes = [C.POINTER(Params), C.c_double]
        L.or_params_size.restype = C.c_size_t
        L.or_setup_jz.argtypes = [C.c_float] * 7 + [_fp, _ip]
        L.or_setup_rp.argtypes = [C.c_int] * 3 + [C.c_float] * 3 + [_fp] * 10
        L.or_load_specs.argtypes = [C.c_char_p, _fp]
        L.or_load_pgm_codes.argtypes = [C.c_char_p, C.c_int, _ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_update_h_f32.argtypes = [C.POINTER(Grid)] + [_fp] * 6
        L.or_update_d_f32.argtypes = [C.POINTER(Grid)] + [_fp] * 9
        L.or_update_sources_f32.argtypes = [C.POINTER(Grid), _fp, _fp, _fp, _fp, C.c_int, _ip, C.c_float]
        L.or_update_e_f32.argtypes = [C.POINTER(Grid), C.c_int] + [_fp] * 22
        L.or_update_mur_f32.argtypes = [C.POINTER(Grid)] + [_fp] * 12
        L.or_sim_create.restype = C.c_void_p
        L.or_sim_create.argtypes = [C.c_int] * 3 + [C.c_float] * 6 + [C.c_int]
        L.or_sim_destroy.argtypes = [C.c_void_p]
        L.or_sim_set_modes.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.or_sim_set_materials.argtypes = [C.c_void_p] + [_fp] * 4
        L.or_sim_set_rp.argtypes = [C.c_void_p, C.c_int, _fp]
        L.or_sim_set_sources.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int, _fp, C.c_int]
        L.or_sim_step.argtypes = [C.c_void_p]
        L.or_sim_run.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _dp]
        L.or_sim_get.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.or_sim_get_f32.argtypes = [C.c_void_p, C.c_int, C.c_int, _fp]
        L.or_sim_set.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.or_sim_set_f32.argtypes = [C.c_void_p, C.c_int, C.c_int, _fp]
        L.or_sim_energy.argtypes = [C.c_void_p]
        L.or_sim_energy.restype = C.c_double
        L.or_sim_steps_done.argtypes = [C.c_void_p]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        assert L.or_params_size() == C.sizeof(Params)
        _lib = L
    return _lib


def set_threads(n):
    lib().or_set_threads(int(n))


def parse_params(path):
    p = Params()
    rc = lib().or_parse_params(path.encode(), C.byref(p))
    if rc != 0:
        raise OSError(f"or_parse_params({path}) -> {rc}")
    return p


def derive(p, timeskip=1.0):
    lib().or_derive_params(C.byref(p), float(timeskip))
    return p


def setup_jz(pulse_width, freq, pmf, dt, dx, dy, dz):
    jz = np.zeros(1 << 16, np.float32)
    info = np.zeros(4, np.int32)
    rc = lib().or_setup_jz(pulse_width, freq, pmf, dt, dx, dy, dz, jz, info)
    if rc != 0:
        raise ValueError(f"or_setup_jz -> {rc}")
    return jz, dict(fine=int(info[0]), temp=int(info[1]), istart=int(info[2]), nraw=int(info[3]))


def setup_rp(shape, freq, pi, eps0, sigma, epsS, epsI, tau):
    nx, ny, nz = shape
    out = [np.zeros(n, np.float32) for n in (2 * ny * nz, nx * 2 * nz, nx * ny * 2) * 2]
    f32 = lambda a: np.ascontiguousarray(a, np.float32).ravel()
    lib().or_setup_rp(nx, ny, nz, freq, pi, eps0, f32(sigma), f32(epsS), f32(epsI), f32(tau), *out)
    return out  # rpx0, rpy0, rpz0, rpxEnd, rpyEnd, rpzEnd


def load_specs(path):
    specs = np.zeros(95 * 4, np.float32)
    rows = lib().or_load_specs(path.encode(), specs)
    if rows < 0:
        raise OSError(f"or_load_specs({path}) -> {rows}")
    return specs.reshape(95, 4), rows


def load_pgm_codes(path, n):
    codes = np.zeros(n, np.int32)
    w, h = C.c_int(), C.c_int()
    rc = lib().or_load_pgm_codes(path.encode(), n, codes, C.byref(w), C.byref(h))
    if rc != 0:
        raise OSError(f"or_load_pgm_codes({path}) -> {rc}")
    return codes, w.value, h.value


def grid(nx, ny, nz, dt, dx, dy, dz, mu0, eps0):
    return Grid(nx, ny, nz, dt, dx, dy, dz, mu0, eps0)


class Sim:
    """A whole simulation: ring slots, materials, rp planes, sources.

    `prec` 32 runs the FP32 transcription, 64 the FP64 oracle.  Arrays are
    numpy with shape (nz, ny, nx) (x fastest, as OFFSET in utils.h:15)."""

    def __init__(self, shape, dt, dx, dy, dz, mu0, eps0, prec=32, debye=DEBYE_FIXED, mur=True):
        self.shape = tuple(shape)  # (nx, ny, nz)
        nx, ny, nz = self.shape
        self.prec = prec
        self._h = lib().or_sim_create(nx, ny, nz, dt, dx, dy, dz, mu0, eps0, prec)
        lib().or_sim_set_modes(self._h, debye, 1 if mur else 0)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_sim_destroy(self._h)
            self._h = None

    @property
    def n(self):
        return lib().or_sim_steps_done(self._h)

    def set_materials(self, sigma, epsS, epsI, tau):
        f = lambda a: np.ascontiguousarray(a, np.float32).ravel()
        lib().or_sim_set_materials(self._h, f(sigma), f(epsS), f(epsI), f(tau))

    def set_rp(self, rps):
        for q, a in enumerate(rps):
            lib().or_sim_set_rp(self._h, q, np.ascontiguousarray(a, np.float32).ravel())

    def set_sources(self, xyz, jz, jz_mode=1):
        xyz = np.ascontiguousarray(np.asarray(xyz, np.int32).reshape(-1, 3))
        jz = np.ascontiguousarray(jz, np.float32)
        lib().or_sim_set_sources(self._h, xyz.shape[0], xyz.ravel(), jz.size, jz, jz_mode)

    def step(self, k=1):
        for _ in range(k):
            lib().or_sim_step(self._h)

    def run(self, steps, probes=()):
        pr = np.ascontiguousarray(np.asarray(probes, np.int32).reshape(-1, 4))
        out = np.zeros((steps, max(pr.shape[0], 1)), np.float64)
        lib().or_sim_run(self._h, steps, pr.shape[0], pr.ravel() if pr.size else np.zeros(4, np.int32), out)
        return out[:, :pr.shape[0]]

    def get(self, field, level=0):
        nx, ny, nz = self.shape
        fid = FIELDS.index(field) if isinstance(field, str) else field
        if self.prec == 32:
            out = np.zeros(nx * ny * nz, np.float32)
            lib().or_sim_get_f32(self._h, fid, level, out)
        else:
            out = np.zeros(nx * ny * nz, np.float64)
            lib().or_sim_get(self._h, fid, level, out)
        return out.reshape(nz, ny, nx)

    def set(self, field, level, arr):
        fid = FIELDS.index(field) if isinstance(field, str) else field
        if self.prec == 32:
            lib().or_sim_set_f32(self._h, fid, level, np.ascontiguousarray(arr, np.float32).ravel())
        else:
            lib().or_sim_set(self._h, fid, level, np.ascontiguousarray(arr, np.float64).ravel())

    def energy(self):
        return lib().or_sim_energy(self._h)
