"""fdtd_amd -- MI355X-native FDTD time-stepping path (Yee H, D from curl H,
Debye E, first-order Mur, soft Jz sources, probes and z-slices).

Python binding of libfdtd_amd.so, whose C ABI is declared in include/:
  * fdtd_calculations.h -- the reference's device entry points
    (copySymbolsToDevice, updateHField, updateDField, updateEField,
    updateSources, updateMurBoundary), fdtd_calculations.h:18-42;
  * fdtd.h -- FdtdParams / FdtdField and the host services
    (initParamsWithPath, setupSources, setupMurBoundary, loadMaterials, ...);
  * fdtd_amd.h -- the context API with the fused gfx950 step, batches of
    grids stepped together, and the serving loop.

The library must be built (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C fdtd_amd/csrc``); there is no CPU fallback: every compute call
goes to the HIP kernels and fails loudly when the library or a GPU is missing.
"""
import ctypes as C
import os
import warnings

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfdtd_amd.so")
CLI_PATH = os.path.join(_HERE, "fdtd")

DEBYE_REF = 0
DEBYE_FIXED = 1
JZ_HOLD3 = 0
JZ_STEP = 1
STORE_F32 = 0
STORE_F16 = 1
STORE_BF16 = 2
STORE_F32_FAST = 3
FIELDS = ["Dx", "Dy", "Dz", "Hx", "Hy", "Hz", "Ex", "Ey", "Ez"]

# Every symbol the headers in include/ declare.
EXPORTS = {
    "fdtd_calculations.h": ["copySymbolsToDevice", "updateHField", "updateDField", "updateEField",
                            "updateSources", "updateMurBoundary", "fdtdSetDebyeMode"],
    "fdtd.h": ["initParamsWithPath", "deallocParams", "printParams", "initFieldWithParams", "deallocField",
               "initDeviceFieldWithParams", "deallocDeviceField", "setupMurBoundary", "loadMaterials",
               "setupSources", "copyDataToDevice", "copyResultsWithParams", "writeResultsWithParams",
               "writeResults", "writeResults3d"],
    "fdtd_amd.h": ["fdtd_last_error", "fdtd_ext_defaults", "fdtd_ctx_create", "fdtd_ctx_destroy",
                   "fdtd_ctx_set_materials", "fdtd_ctx_set_material_arrays", "fdtd_ctx_set_rp",
                   "fdtd_ctx_set_sources", "fdtd_ctx_set_probes", "fdtd_ctx_set_slices", "fdtd_ctx_set_field",
                   "fdtd_ctx_get_field", "fdtd_ctx_field_ptr", "fdtd_step_fused", "fdtd_run",
                   "fdtd_ctx_steps_done", "fdtd_ctx_snapshot", "fdtd_probe_read", "fdtd_probe_peak_units", "fdtd_slices_read",
                   "fdtd_ctx_launch_shape", "fdtd_default_launch_shape",
                   "fdtd_ctx_set_timing", "fdtd_ctx_kernel_time", "fdtd_ctx_timed_steps", "fdtd_ctx_scales", "fdtd_ctx_layout",
                   "fdtd_ctx_set_energy", "fdtd_energy_read", "fdtd_ctx_save", "fdtd_ctx_load",
                   "fdtd_ctx_autotune", "fdtd_ctx_halo_bytes", "fdtd_ctx_halo_pack", "fdtd_ctx_halo_unpack",
                   "fdtd_batch_create", "fdtd_batch_destroy", "fdtd_batch_instance_bytes", "fdtd_batch_add",
                   "fdtd_batch_remove", "fdtd_batch_run", "fdtd_batch_status", "fdtd_batch_probe_read",
                   "fdtd_batch_probe_peak_units",
                   "fdtd_batch_get_field", "fdtd_batch_set_shape", "fdtd_batch_autotune", "fdtd_batch_info",
                   "fdtd_dropin_solve", "fdtd_dropin_solve_noalloc", "fdtd_stream_probe", "fdtd_stream_probe_layout", "fdtd_load_tissue_table", "fdtd_load_phantom_ids",
                   "fdtd_serve_create", "fdtd_serve_destroy", "fdtd_serve_submit", "fdtd_serve_wait",
                   "fdtd_serve_result", "fdtd_serve_release", "fdtd_serve_drain", "fdtd_serve_stats"],
}
EINVAL, ERANGE, ESTATE, EIO, ENOMEM = -1, -2, -3, -4, -5
PGM_REFERENCE, PGM_IMAGE = 0, 1
HALO_UP, HALO_DOWN = 0, 1
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class FdtdParams(C.Structure):
    """fdtd.h:7-34, field for field."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("iterationsCount", C.c_int),
        ("inputPath", C.c_char_p), ("outputPath", C.c_char_p),
        ("elementsPerWave", C.c_int),
        ("waveFrequency", C.c_float), ("pulseWidth", C.c_float), ("pulseModulationFrequency", C.c_float),
        ("sourcesCount", C.c_int), ("sources", C.POINTER(C.c_int)),
        ("defaultSigma", C.c_float), ("defaultEpsS", C.c_float), ("defaultEpsI", C.c_float),
        ("defaultTauD", C.c_float),
        ("jzCount", C.c_int), ("jz", C.POINTER(C.c_float)),
        ("pi", C.c_float), ("c", C.c_float), ("timeskip", C.c_float), ("lambda_", C.c_float),
        ("dt", C.c_float), ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("mu0", C.c_float), ("eps0", C.c_float),
    ]


_FP = C.POINTER(C.c_float)


class FdtdField(C.Structure):
    """fdtd.h:37-54."""
    _fields_ = [(n, _FP) for n in (
        "ex0", "ey0", "ez0", "ex1", "ey1", "ez1", "ex2", "ey2", "ez2", "hx", "hy", "hz",
        "dx0", "dy0", "dz0", "dx1", "dy1", "dz1", "dx2", "dy2", "dz2",
        "epsI", "epsS", "tauD", "sigma", "rpx0", "rpy0", "rpz0", "rpxEnd", "rpyEnd", "rpzEnd")]


class FdtdExt(C.Structure):
    _fields_ = [("debye_mode", C.c_int), ("jz_index_mode", C.c_int), ("jz_limit", C.c_int),
                ("tile_y", C.c_int), ("kchunk", C.c_int), ("xcd_remap", C.c_int), ("storage", C.c_int),
                ("scale_e", C.c_float), ("scale_d", C.c_float), ("scale_h", C.c_float), ("lane_cells", C.c_int),
                ("tile_x", C.c_int), ("load_zero_terms", C.c_int), ("time_steps", C.c_int)]


class FdtdInstance(C.Structure):
    """One grid of a batch or a served request (fdtd_amd.h)."""
    _fields_ = [("ids", C.c_void_p), ("nmat", C.c_int), ("table", C.c_void_p), ("rp", C.c_void_p * 6),
                ("nsrc", C.c_int), ("src_xyz", C.c_void_p), ("jz_len", C.c_int), ("jz", C.c_void_p),
                ("steps", C.c_int), ("nprobe", C.c_int), ("probes", C.c_void_p)]


class FdtdBatchInfo(C.Structure):
    _fields_ = [("instances", C.c_int), ("running", C.c_int), ("step", C.c_longlong),
                ("arena_bytes", C.c_size_t), ("used_bytes", C.c_size_t), ("largest_free", C.c_size_t),
                ("tile_y", C.c_int), ("kchunk", C.c_int), ("graph_steps", C.c_int),
                ("grid", C.c_int * 4), ("tiles", C.c_int * 4), ("cells", C.c_longlong * 4)]


class FdtdServeConfig(C.Structure):
    _fields_ = [("hbm_bytes", C.c_size_t), ("max_slots", C.c_int), ("segment_steps", C.c_int),
                ("tile_y", C.c_int), ("kchunk", C.c_int)]


class FdtdRequestInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("error", C.c_int), ("rows", C.c_int), ("nprobe", C.c_int),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("cells", C.c_longlong), ("steps", C.c_int)]


class FdtdServeStats(C.Structure):
    _fields_ = [("submitted", C.c_int), ("completed", C.c_int), ("failed", C.c_int), ("queued", C.c_int),
                ("running", C.c_int), ("t_first_submit", C.c_double), ("t_last_done", C.c_double),
                ("solves_per_s", C.c_double), ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double),
                ("mean_latency_s", C.c_double), ("cell_updates", C.c_double), ("segments", C.c_longlong),
                ("peak_bytes", C.c_size_t), ("peak_running", C.c_int)]


class FdtdError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libfdtd_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FdtdError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                            "or `make -C fdtd_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        vp, ip, i = C.c_void_p, C.c_int, C.c_int
        L.fdtd_last_error.restype = C.c_char_p
        L.fdtd_ext_defaults.argtypes = [C.POINTER(FdtdExt)]
        L.fdtd_ctx_create.argtypes = [C.POINTER(FdtdParams), C.POINTER(FdtdExt), C.POINTER(vp)]
        L.fdtd_ctx_destroy.argtypes = [vp]
        L.fdtd_ctx_set_materials.argtypes = [vp, vp, i, vp]
        L.fdtd_ctx_set_material_arrays.argtypes = [vp, vp, vp, vp, vp]
        L.fdtd_ctx_set_rp.argtypes = [vp] * 7
        L.fdtd_ctx_set_sources.argtypes = [vp, i, vp, i, vp]
        L.fdtd_ctx_set_probes.argtypes = [vp, i, vp, i]
        L.fdtd_ctx_set_slices.argtypes = [vp, i, i, i]
        L.fdtd_ctx_set_field.argtypes = [vp, i, i, vp]
        L.fdtd_ctx_get_field.argtypes = [vp, i, i, vp]
        L.fdtd_ctx_field_ptr.argtypes = [vp, i, i, C.POINTER(vp)]
        L.fdtd_step_fused.argtypes = [vp, vp]
        L.fdtd_run.argtypes = [vp, i, vp]
        L.fdtd_ctx_steps_done.argtypes = [vp]
        L.fdtd_ctx_snapshot.argtypes = [vp, vp, vp]
        L.fdtd_probe_read.argtypes = [vp, vp, i]
        L.fdtd_probe_peak_units.argtypes = [vp, vp, i]
        L.fdtd_slices_read.argtypes = [vp, vp, vp, i]
        L.fdtd_ctx_launch_shape.argtypes = [vp, vp, vp, vp, vp]
        L.fdtd_default_launch_shape.argtypes = [C.POINTER(FdtdParams), C.POINTER(FdtdExt), i, vp, vp, vp, vp]
        L.fdtd_ctx_set_timing.argtypes = [vp, i, i]
        L.fdtd_ctx_kernel_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.fdtd_ctx_timed_steps.argtypes = [vp]
        L.fdtd_ctx_scales.argtypes = [vp] + [C.POINTER(C.c_double)] * 3
        L.fdtd_ctx_set_energy.argtypes = [vp, i, i]
        L.fdtd_ctx_save.argtypes = [vp, C.c_char_p]
        L.fdtd_ctx_autotune.argtypes = [vp, i, vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.fdtd_ctx_load.argtypes = [vp, C.c_char_p]
        L.fdtd_energy_read.argtypes = [vp, vp, vp, i]
        L.fdtd_ctx_layout.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
        L.fdtd_ctx_halo_bytes.argtypes = [vp, i, C.POINTER(C.c_size_t)]
        L.fdtd_ctx_halo_pack.argtypes = [vp, i, vp, vp]
        L.fdtd_ctx_halo_unpack.argtypes = [vp, i, vp, vp]
        PP, PE, PI = C.POINTER(FdtdParams), C.POINTER(FdtdExt), C.POINTER(FdtdInstance)
        L.fdtd_batch_create.argtypes = [C.c_size_t, C.POINTER(vp)]
        L.fdtd_batch_destroy.argtypes = [vp]
        L.fdtd_batch_instance_bytes.argtypes = [PP, PE, PI, C.POINTER(C.c_size_t)]
        L.fdtd_batch_add.argtypes = [vp, PP, PE, PI, C.POINTER(C.c_int)]
        L.fdtd_batch_remove.argtypes = [vp, i]
        L.fdtd_batch_run.argtypes = [vp, i, vp]
        L.fdtd_batch_status.argtypes = [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.fdtd_batch_probe_read.argtypes = [vp, i, vp, i]
        L.fdtd_batch_probe_peak_units.argtypes = [vp, i, vp, i]
        L.fdtd_batch_get_field.argtypes = [vp, i, i, i, vp]
        L.fdtd_batch_set_shape.argtypes = [vp, i, i, i]
        L.fdtd_batch_autotune.argtypes = [vp, i, vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.fdtd_batch_info.argtypes = [vp, C.POINTER(FdtdBatchInfo)]
        L.fdtd_dropin_solve.argtypes = [PP, PE, PI, vp, vp]
        L.fdtd_dropin_solve_noalloc.argtypes = [PP, PE, PI, vp, vp]
        L.fdtd_stream_probe.argtypes = [C.c_longlong, i, i, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.fdtd_stream_probe_layout.argtypes = [C.c_longlong, i, i, i, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.fdtd_load_tissue_table.argtypes = [C.c_char_p, vp]
        L.fdtd_load_phantom_ids.argtypes = [C.c_char_p, i, i, i, i, i, vp]
        L.fdtd_serve_create.argtypes = [C.POINTER(FdtdServeConfig), C.POINTER(vp)]
        L.fdtd_serve_destroy.argtypes = [vp]
        L.fdtd_serve_submit.argtypes = [vp, PP, PE, PI, C.POINTER(C.c_longlong)]
        L.fdtd_serve_wait.argtypes = [vp, C.c_longlong, C.c_double]
        L.fdtd_serve_result.argtypes = [vp, C.c_longlong, vp, i, C.POINTER(FdtdRequestInfo)]
        L.fdtd_serve_release.argtypes = [vp, C.c_longlong]
        L.fdtd_serve_drain.argtypes = [vp, C.c_double]
        L.fdtd_serve_stats.argtypes = [vp, C.POINTER(FdtdServeStats)]
        L.copySymbolsToDevice.argtypes = [C.POINTER(FdtdParams)]
        L.updateHField.argtypes = [vp] * 7
        L.updateDField.argtypes = [vp] * 10
        L.updateEField.argtypes = [vp] * 23
        L.updateSources.argtypes = [vp, vp, vp, vp, i, vp]
        L.updateMurBoundary.argtypes = [vp] * 13
        L.fdtdSetDebyeMode.argtypes = [i]
        L.initParamsWithPath.argtypes = [C.c_char_p]
        L.initParamsWithPath.restype = C.POINTER(FdtdParams)
        L.deallocParams.argtypes = [C.POINTER(FdtdParams)]
        L.setupSources.argtypes = [C.POINTER(FdtdParams)]
        L.initFieldWithParams.argtypes = [C.POINTER(FdtdParams)]
        L.initFieldWithParams.restype = C.POINTER(FdtdField)
        L.deallocField.argtypes = [C.POINTER(FdtdField)]
        L.setupMurBoundary.argtypes = [C.POINTER(FdtdParams), C.POINTER(FdtdField)]
        L.loadMaterials.argtypes = [C.POINTER(FdtdParams), C.POINTER(FdtdField), C.c_char_p, C.c_char_p]
        L.writeResults.argtypes = [C.POINTER(FdtdParams)] + [vp] * 9 + [i]
        L.writeResults3d.argtypes = [C.POINTER(FdtdParams)] + [vp] * 9 + [i]
        _lib = L
    return _lib


class PrecisionWarning(UserWarning):
    """A 16-bit result lost the precision of its storage format."""


def _warn_low_probes(peak, low, where=""):
    if low > 0:
        idx = [int(p) for p in np.nonzero((peak > 0) & (peak < 2.0 ** -14))[0]]
        warnings.warn(f"{where}fp16 storage: probe(s) {idx} peaked below fp16's normal range (2^-14 stored units; "
                      f"peaks {[float(f'{peak[p]:.3g}') for p in idx]}): their fields never had fp16's 11 "
                      f"significant bits and may be far off; one scale per field cannot hold the sources' near "
                      f"field and these probes at once -- use fp32 storage", PrecisionWarning, stacklevel=3)


def check(rc, what=""):
    if rc != 0:
        msg = lib().fdtd_last_error().decode(errors="replace")
        raise FdtdError(f"{what} failed with code {rc}: {msg}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_params(shape, dt, dx, dy=None, dz=None, mu0=None, eps0=None):
    """An FdtdParams with the scalars the device path uses (others zero)."""
    nx, ny, nz = shape
    p = FdtdParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.dt = dt
    p.dx = dx
    p.dy = dx if dy is None else dy
    p.dz = dx if dz is None else dz
    p.mu0 = mu0
    p.eps0 = eps0
    return p


class Context:
    """One grid's device state and the fused step (fdtd_amd.h)."""

    def __init__(self, params, debye=DEBYE_FIXED, jz_mode=JZ_STEP, jz_limit=0, tile_y=0, kchunk=0, xcd_remap=None,
                 storage=STORE_F32, scales=(0.0, 0.0, 0.0), lane_cells=0, tile_x=0, load_zero_terms=0,
                 time_steps=0):
        L = lib()
        ext = FdtdExt()
        L.fdtd_ext_defaults(C.byref(ext))
        ext.debye_mode, ext.jz_index_mode, ext.jz_limit = debye, jz_mode, jz_limit
        ext.tile_y, ext.kchunk = tile_y, kchunk
        if xcd_remap is not None:      # else the library's default order
            ext.xcd_remap = xcd_remap
        ext.storage = storage
        ext.scale_e, ext.scale_d, ext.scale_h = scales
        ext.lane_cells = lane_cells
        ext.tile_x = tile_x
        ext.load_zero_terms = load_zero_terms
        ext.time_steps = time_steps
        h = C.c_void_p()
        self._params = params
        check(L.fdtd_ctx_create(C.byref(params), C.byref(ext), C.byref(h)), "fdtd_ctx_create")
        self._h = h
        self.shape = (params.nx, params.ny, params.nz)
        self.nprobe = 0
        self.storage = storage

    def close(self):
        if getattr(self, "_h", None):
            lib().fdtd_ctx_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_materials(self, ids, table):
        ids = np.ascontiguousarray(ids, np.uint8).ravel()
        table = np.ascontiguousarray(table, np.float32).reshape(-1, 4)
        check(lib().fdtd_ctx_set_materials(self._h, _ptr(ids), table.shape[0], _ptr(table)), "set_materials")

    def set_material_arrays(self, sigma, epsS, epsI, tau):
        a = [np.ascontiguousarray(x, np.float32).ravel() for x in (sigma, epsS, epsI, tau)]
        check(lib().fdtd_ctx_set_material_arrays(self._h, *[_ptr(x) for x in a]), "set_material_arrays")

    def set_rp(self, rps):
        a = [np.ascontiguousarray(x, np.float32).ravel() for x in rps]
        check(lib().fdtd_ctx_set_rp(self._h, *[_ptr(x) for x in a]), "set_rp")

    def set_sources(self, xyz, jz):
        xyz = np.ascontiguousarray(np.asarray(xyz, np.int32).reshape(-1, 3))
        jz = np.ascontiguousarray(jz, np.float32)
        check(lib().fdtd_ctx_set_sources(self._h, xyz.shape[0], _ptr(xyz), jz.size, _ptr(jz)), "set_sources")

    def set_probes(self, probes, capacity):
        pr = np.ascontiguousarray(np.asarray(probes, np.int32).reshape(-1, 4))
        self.nprobe = pr.shape[0]
        check(lib().fdtd_ctx_set_probes(self._h, pr.shape[0], _ptr(pr), capacity), "set_probes")

    def set_slices(self, k, interval, capacity):
        check(lib().fdtd_ctx_set_slices(self._h, k, interval, capacity), "set_slices")

    def autotune(self, trial_steps=4, stream=None):
        """(tile_y, kchunk, lane_cells) picked by timing each shape; the context returns to step 0."""
        v = [C.c_int() for _ in range(3)]
        check(lib().fdtd_ctx_autotune(self._h, trial_steps, stream, *[C.byref(x) for x in v]), "autotune")
        return tuple(x.value for x in v)

    def save(self, path):
        check(lib().fdtd_ctx_save(self._h, str(path).encode()), "save")

    def load(self, path):
        check(lib().fdtd_ctx_load(self._h, str(path).encode()), "load")

    def set_energy(self, interval, capacity):
        check(lib().fdtd_ctx_set_energy(self._h, interval, capacity), "set_energy")

    def energy(self, max_rows=100000):
        """(step counts, fp64 energy sums) recorded so far."""
        steps = np.zeros(max_rows, np.int32)
        e = np.zeros(max_rows, np.float64)
        got = lib().fdtd_energy_read(self._h, _ptr(steps), _ptr(e), max_rows)
        if got < 0:
            check(got, "energy_read")
        return steps[:got], e[:got]

    def set(self, field, level, arr):
        fid = FIELDS.index(field) if isinstance(field, str) else field
        a = np.ascontiguousarray(arr, np.float32).ravel()
        check(lib().fdtd_ctx_set_field(self._h, fid, level, _ptr(a)), "set_field")

    def get(self, field, level=0):
        fid = FIELDS.index(field) if isinstance(field, str) else field
        nx, ny, nz = self.shape
        out = np.empty(nx * ny * nz, np.float32)
        check(lib().fdtd_ctx_get_field(self._h, fid, level, _ptr(out)), "get_field")
        return out.reshape(nz, ny, nx)

    def field_ptr(self, field, level=0):
        fid = FIELDS.index(field) if isinstance(field, str) else field
        p = C.c_void_p()
        check(lib().fdtd_ctx_field_ptr(self._h, fid, level, C.byref(p)), "field_ptr")
        return p.value

    def step(self, stream=None):
        check(lib().fdtd_step_fused(self._h, stream), "fdtd_step_fused")

    def run(self, steps, stream=None):
        check(lib().fdtd_run(self._h, steps, stream), "fdtd_run")

    @property
    def n(self):
        return lib().fdtd_ctx_steps_done(self._h)

    def probes(self, max_rows=None):
        """Recorded probe rows (SI units).  Under fp16 storage, warns when a
        probe has so far peaked below fp16's normal range (probe_peak_units)."""
        rows = self.n if max_rows is None else max_rows
        out = np.zeros((max(rows, 1), max(self.nprobe, 1)), np.float64)
        got = lib().fdtd_probe_read(self._h, _ptr(out), rows)
        if got < 0:
            check(got, "probe_read")
        if self.storage == STORE_F16 and self.nprobe and got > 0:
            _warn_low_probes(*self.probe_peak_units())
        return out[:got, :self.nprobe]

    def probe_peak_units(self):
        """(peak |value| of each probe so far in stored units, how many saw a
        signal that peaked below 2^-14, fp16's smallest normal number, under
        fp16 storage)."""
        peak = np.zeros(max(self.nprobe, 1), np.float64)
        low = lib().fdtd_probe_peak_units(self._h, _ptr(peak), self.nprobe)
        if low < 0:
            check(low, "probe_peak_units")
        return peak[:self.nprobe], low

    def snapshot(self, dev_out, stream=None):
        """Enqueue the nine newest fields, packed fp32 in SI units, into the
        device buffer at address dev_out (9*nx*ny*nz floats)."""
        check(lib().fdtd_ctx_snapshot(self._h, dev_out, stream), "snapshot")

    def slices(self, max_slices=1024):
        nx, ny, _ = self.shape
        steps = np.zeros(max_slices, np.int32)
        data = np.zeros((max_slices, 9, ny, nx), np.float32)
        got = lib().fdtd_slices_read(self._h, _ptr(steps), _ptr(data), max_slices)
        if got < 0:
            check(got, "slices_read")
        return steps[:got], data[:got]

    def set_timing(self, capacity, every=1):
        """Time every `every`-th fused launch (the first included), at most
        `capacity` of them."""
        check(lib().fdtd_ctx_set_timing(self._h, capacity, every), "set_timing")

    def kernel_time(self):
        """(summed fused-kernel milliseconds, launches timed)."""
        ms, n = C.c_double(), C.c_int()
        check(lib().fdtd_ctx_kernel_time(self._h, C.byref(ms), C.byref(n)), "kernel_time")
        return ms.value, n.value

    def timed_steps(self):
        """Steps per timed launch: 1, or 2 for two-step launches (0: none timed)."""
        return lib().fdtd_ctx_timed_steps(self._h)

    def scales(self):
        v = [C.c_double() for _ in range(3)]
        check(lib().fdtd_ctx_scales(self._h, *[C.byref(x) for x in v]), "scales")
        return tuple(x.value for x in v)

    def layout(self):
        """(pitch, plane, element bytes) of the device fields."""
        p, pl, e = C.c_int(), C.c_longlong(), C.c_int()
        check(lib().fdtd_ctx_layout(self._h, C.byref(p), C.byref(pl), C.byref(e)), "layout")
        return p.value, pl.value, e.value

    def halo_bytes(self, direction):
        """Bytes of a z-slab halo message (HALO_UP or HALO_DOWN)."""
        n = C.c_size_t()
        check(lib().fdtd_ctx_halo_bytes(self._h, direction, C.byref(n)), "halo_bytes")
        return n.value

    def halo_pack(self, direction, dst, stream=None):
        """Pack this slab's outgoing halo planes into device buffer `dst`."""
        check(lib().fdtd_ctx_halo_pack(self._h, direction, dst, stream), "halo_pack")

    def halo_unpack(self, direction, src, stream=None):
        """Unpack a neighbour's halo message (device buffer) into a ghost plane."""
        check(lib().fdtd_ctx_halo_unpack(self._h, direction, src, stream), "halo_unpack")

    def launch_shape(self):
        v = [C.c_int() for _ in range(4)]
        check(lib().fdtd_ctx_launch_shape(self._h, *[C.byref(x) for x in v]), "launch_shape")
        return dict(tile_x=v[0].value, tile_y=v[1].value, kchunk=v[2].value, blocks=v[3].value)


def make_ext(debye=DEBYE_FIXED, jz_mode=JZ_STEP, jz_limit=0, storage=STORE_F32, scales=(0.0, 0.0, 0.0),
             tile_y=0, kchunk=0, xcd_remap=None, load_zero_terms=0, time_steps=0):
    ext = FdtdExt()
    lib().fdtd_ext_defaults(C.byref(ext))
    ext.debye_mode, ext.jz_index_mode, ext.jz_limit = debye, jz_mode, jz_limit
    ext.tile_y, ext.kchunk = tile_y, kchunk
    if xcd_remap is not None:
        ext.xcd_remap = xcd_remap
    ext.storage = storage
    ext.scale_e, ext.scale_d, ext.scale_h = scales
    ext.load_zero_terms = load_zero_terms
    ext.time_steps = time_steps
    return ext


def default_launch_shape(params, storage=STORE_F32, cus=0, **ext_fields):
    """The launch shape a Context would take for this grid (no device memory
    is touched): dict(tile_x, tile_y, kchunk, lane_cells).  cus = compute
    units of the target GPU (0: the current device's, 256 without one)."""
    ext = make_ext(storage=storage)
    for k, v in ext_fields.items():
        setattr(ext, k, v)
    v = [C.c_int() for _ in range(4)]
    check(lib().fdtd_default_launch_shape(C.byref(params), C.byref(ext), cus, *[C.byref(x) for x in v]),
          "fdtd_default_launch_shape")
    return dict(tile_x=v[0].value, tile_y=v[1].value, kchunk=v[2].value, lane_cells=v[3].value)


class Instance:
    """Inputs of one batched or served grid; keeps the arrays alive while the
    FdtdInstance that points at them is in use."""

    def __init__(self, params, ext, ids, table, src_xyz, jz, steps, probes=(), rp=None):
        self.params, self.ext = params, ext
        self.shape = (params.nx, params.ny, params.nz)
        self.ids = None if ids is None else np.ascontiguousarray(ids, np.uint8).ravel()
        self.table = np.ascontiguousarray(table, np.float32).reshape(-1, 4)
        self.src = np.ascontiguousarray(np.asarray(src_xyz, np.int32).reshape(-1, 3))
        self.jz = np.ascontiguousarray(jz, np.float32)
        self.probes = np.ascontiguousarray(np.asarray(probes, np.int32).reshape(-1, 4))
        self.rp = None if rp is None else [np.ascontiguousarray(x, np.float32).ravel() for x in rp]
        self.steps = int(steps)
        s = FdtdInstance()
        s.ids = _ptr(self.ids)
        s.nmat = self.table.shape[0]
        s.table = _ptr(self.table)
        for r in range(6):
            s.rp[r] = _ptr(self.rp[r]) if self.rp is not None else None
        s.nsrc = self.src.shape[0]
        s.src_xyz = _ptr(self.src)
        s.jz_len = self.jz.size
        s.jz = _ptr(self.jz)
        s.steps = self.steps
        s.nprobe = self.probes.shape[0]
        s.probes = _ptr(self.probes)
        self.c = s

    @property
    def nprobe(self):
        return self.probes.shape[0]

    def nbytes(self):
        b = C.c_size_t()
        check(lib().fdtd_batch_instance_bytes(C.byref(self.params), C.byref(self.ext), C.byref(self.c),
                                              C.byref(b)), "instance_bytes")
        return b.value


def stream_probe(cells, blocks=0, reps=20):
    """(GB/s, ms per pass) of the fused step's 15-read / 9-write access mix
    streamed without arithmetic over `cells` floats per array
    (fdtd_stream_probe); blocks 0 = as many as fit resident."""
    g, ms = C.c_double(), C.c_double()
    check(lib().fdtd_stream_probe(int(cells), int(blocks), int(reps), C.byref(g), C.byref(ms)), "fdtd_stream_probe")
    return g.value, ms.value


def stream_probe_layout(cells, row, blocks=0, reps=20):
    """(GB/s, ms per pass) of the same mix with the 24 arrays interleaved by
    rows of `row` elements (row records; 0 = separate arrays)."""
    g, ms = C.c_double(), C.c_double()
    check(lib().fdtd_stream_probe_layout(int(cells), int(row), int(blocks), int(reps), C.byref(g), C.byref(ms)),
          "fdtd_stream_probe_layout")
    return g.value, ms.value


def load_tissue_table(path):
    """95 x {sigma, eps_s, eps_inf, tau} rows (row = tissue code) and the rows read."""
    t = np.zeros((95, 4), np.float32)
    rows = lib().fdtd_load_tissue_table(str(path).encode(), _ptr(t))
    if rows < 0:
        check(rows, "load_tissue_table")
    return t, rows


def load_phantom_ids(directory, shape, mode=PGM_REFERENCE, threads=0):
    """uint8 tissue codes (nz, ny, nx) of <directory>/v1_%05d.pgm."""
    nx, ny, nz = shape
    ids = np.zeros((nz, ny, nx), np.uint8)
    rc = lib().fdtd_load_phantom_ids(str(directory).encode(), nx, ny, nz, mode, threads, _ptr(ids))
    if rc < 0:
        check(rc, "load_phantom_ids")
    return ids


def dropin_solve(inst, stream=None):
    """Probe series [steps, nprobe] of one grid solved through the five
    drop-in entry points (fdtd_dropin_solve)."""
    out = np.zeros((inst.steps, max(inst.nprobe, 1)), np.float64)
    check(lib().fdtd_dropin_solve(C.byref(inst.params), C.byref(inst.ext), C.byref(inst.c), _ptr(out), stream),
          "fdtd_dropin_solve")
    return out[:, :inst.nprobe]


def dropin_solve_noalloc(inst, stream=None):
    """As dropin_solve, with the grid's state in the library's static device
    array instead of allocations (fdtd_dropin_solve_noalloc)."""
    out = np.zeros((inst.steps, max(inst.nprobe, 1)), np.float64)
    check(lib().fdtd_dropin_solve_noalloc(C.byref(inst.params), C.byref(inst.ext), C.byref(inst.c), _ptr(out),
                                          stream), "fdtd_dropin_solve_noalloc")
    return out[:, :inst.nprobe]


class Batch:
    """Many grids stepped together (fdtd_batch_*)."""

    def __init__(self, hbm_bytes):
        h = C.c_void_p()
        check(lib().fdtd_batch_create(hbm_bytes, C.byref(h)), "fdtd_batch_create")
        self._h = h
        self._inst = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().fdtd_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, inst):
        i = C.c_int()
        check(lib().fdtd_batch_add(self._h, C.byref(inst.params), C.byref(inst.ext), C.byref(inst.c),
                                   C.byref(i)), "fdtd_batch_add")
        self._inst[i.value] = inst
        return i.value

    def remove(self, i):
        check(lib().fdtd_batch_remove(self._h, i), "fdtd_batch_remove")
        self._inst.pop(i, None)

    def run(self, steps, stream=None):
        check(lib().fdtd_batch_run(self._h, steps, stream), "fdtd_batch_run")

    def status(self, i):
        d, t = C.c_int(), C.c_int()
        check(lib().fdtd_batch_status(self._h, i, C.byref(d), C.byref(t)), "fdtd_batch_status")
        return d.value, t.value

    def probes(self, i):
        """Grid i's recorded probe rows (SI units); warns as Context.probes
        does under fp16 storage."""
        inst = self._inst[i]
        rows = inst.steps
        out = np.zeros((max(rows, 1), max(inst.nprobe, 1)), np.float64)
        got = lib().fdtd_batch_probe_read(self._h, i, _ptr(out), rows)
        if got < 0:
            check(got, "fdtd_batch_probe_read")
        if inst.ext.storage == STORE_F16 and inst.nprobe and got > 0:
            _warn_low_probes(*self.probe_peak_units(i), f"batch grid {i}: ")
        return out[:got, :inst.nprobe]

    def probe_peak_units(self, i):
        """Context.probe_peak_units for grid i."""
        n = self._inst[i].nprobe
        peak = np.zeros(max(n, 1), np.float64)
        low = lib().fdtd_batch_probe_peak_units(self._h, i, _ptr(peak), n)
        if low < 0:
            check(low, "fdtd_batch_probe_peak_units")
        return peak[:n], low

    def get(self, i, field, level=0):
        fid = FIELDS.index(field) if isinstance(field, str) else field
        nx, ny, nz = self._inst[i].shape
        out = np.empty(nx * ny * nz, np.float32)
        check(lib().fdtd_batch_get_field(self._h, i, fid, level, _ptr(out)), "fdtd_batch_get_field")
        return out.reshape(nz, ny, nx)

    def set_shape(self, tile_y=0, kchunk=0, graph_steps=0):
        check(lib().fdtd_batch_set_shape(self._h, tile_y, kchunk, graph_steps), "fdtd_batch_set_shape")

    def autotune(self, trial_steps=8, stream=None):
        ty, kc = C.c_int(), C.c_int()
        check(lib().fdtd_batch_autotune(self._h, trial_steps, stream, C.byref(ty), C.byref(kc)), "autotune")
        return ty.value, kc.value

    def info(self):
        inf = FdtdBatchInfo()
        check(lib().fdtd_batch_info(self._h, C.byref(inf)), "fdtd_batch_info")
        return {k: (list(getattr(inf, k)) if k in ("grid", "tiles", "cells") else getattr(inf, k))
                for k, _ in FdtdBatchInfo._fields_}


class Server:
    """Continuous-batching request loop (fdtd_serve_*)."""

    def __init__(self, hbm_bytes, max_slots=64, segment_steps=8, tile_y=0, kchunk=0):
        cfg = FdtdServeConfig(hbm_bytes, max_slots, segment_steps, tile_y, kchunk)
        h = C.c_void_p()
        check(lib().fdtd_serve_create(C.byref(cfg), C.byref(h)), "fdtd_serve_create")
        self._h = h
        self._req = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().fdtd_serve_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, inst):
        r = C.c_longlong()
        check(lib().fdtd_serve_submit(self._h, C.byref(inst.params), C.byref(inst.ext), C.byref(inst.c),
                                      C.byref(r)), "fdtd_serve_submit")
        self._req[r.value] = (inst.steps, inst.nprobe)
        return r.value

    def wait(self, r, timeout=-1.0):
        rc = lib().fdtd_serve_wait(self._h, r, timeout)
        if rc < 0:
            check(rc, "fdtd_serve_wait")
        return rc == 0

    def result(self, r):
        """(probe series [steps, nprobe], info dict)."""
        steps, nprobe = self._req[r]
        out = np.zeros((max(steps, 1), max(nprobe, 1)), np.float64)
        info = FdtdRequestInfo()
        got = lib().fdtd_serve_result(self._h, r, _ptr(out), steps, C.byref(info))
        if got < 0:
            check(got, "fdtd_serve_result")
        return out[:got, :nprobe], {k: getattr(info, k) for k, _ in FdtdRequestInfo._fields_}

    def release(self, r):
        check(lib().fdtd_serve_release(self._h, r), "fdtd_serve_release")
        self._req.pop(r, None)

    def drain(self, timeout=-1.0):
        rc = lib().fdtd_serve_drain(self._h, timeout)
        if rc < 0:
            check(rc, "fdtd_serve_drain")
        return rc == 0

    def stats(self):
        st = FdtdServeStats()
        check(lib().fdtd_serve_stats(self._h, C.byref(st)), "fdtd_serve_stats")
        return {k: getattr(st, k) for k, _ in FdtdServeStats._fields_}
