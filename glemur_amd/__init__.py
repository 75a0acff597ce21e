"""Python bindings of libglemur (include/glemur.h) through ctypes.

The library is the product: HIP kernels for gfx950 behind a C ABI.  This
module only marshals arguments; there is no Python or CPU fallback -- if the
shared library is missing, importing the package fails.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libglemur.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libglemur.so not found at {LIB_PATH}; build it first "
        "(python -c 'import __graft_entry__ as g; g.build()')")

_lib = C.CDLL(LIB_PATH)

GM_OK = 0
TIMING_KINDS = 7      # gm_timing_read: stages 0..5, controller
GM_ERR_ARG, GM_ERR_HIP, GM_ERR_IO, GM_ERR_UNSUPPORTED, GM_ERR_NAN, GM_ERR_PARSE = -1, -2, -3, -4, -5, -6
PREC_F64, PREC_F32K = 0, 1
FIELD_XB, FIELD_B0, FIELD_B, FIELD_DETJAC, FIELD_J, FIELD_GRADP, FIELD_KK, FIELD_XB_NEW, FIELD_XB_LO = range(9)
FIELD_CELLVOL, FIELD_CONVEXITY, FIELD_WEDGEMIN = 9, 10, 11
FIELD_UU = 12


class Params(C.Structure):
    """gm_params: the reference's struct Parameters with REAL = double
    (src/global.h:44-108)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("Ox", C.c_double), ("Oy", C.c_double), ("Oz", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dx1", C.c_double), ("dy1", C.c_double), ("dz1", C.c_double),
        ("bInit", C.c_char * 30), ("uInit", C.c_char * 30),
        ("ampl", C.c_double), ("phi1", C.c_double), ("phi2", C.c_double),
        ("L1", C.c_double), ("L2", C.c_double), ("ar", C.c_double), ("az", C.c_double),
        ("nBlobs", C.c_int), ("width", C.c_double), ("minor", C.c_double), ("major", C.c_double),
        ("stretch", C.c_double), ("bGround", C.c_double),
        ("blobXc", C.c_double * 10), ("blobYc", C.c_double * 10), ("blobZc", C.c_double * 10),
        ("blobZl", C.c_double * 10), ("blobTwist", C.c_double * 10), ("blobScale", C.c_double * 10),
        ("initDist", C.c_char * 30), ("initDistCode", C.c_int),
        ("initShearA", C.c_double), ("initShearB", C.c_double), ("initShearK", C.c_double),
        ("pert", C.c_double), ("twist", C.c_double), ("fRestart", C.c_bool),
        ("nCycle", C.c_int), ("dt0", C.c_double), ("dtMin", C.c_double), ("maxError", C.c_double),
        ("zUpdate", C.c_bool), ("inertia", C.c_bool), ("pressure", C.c_bool),
        ("nu", C.c_double), ("beta", C.c_double), ("jxbAver", C.c_bool), ("jxbAverWeight", C.c_double),
        ("xPeri", C.c_bool), ("yPeri", C.c_bool), ("zPeri", C.c_bool), ("epsilonProf", C.c_bool),
        ("jMethod", C.c_char * 30), ("pMethod", C.c_char * 30),
        ("nTs", C.c_int), ("dtDump", C.c_double), ("dtSave", C.c_double),
        ("dumpJ", C.c_bool), ("dumpDetJac", C.c_bool), ("dumpCellVol", C.c_bool),
        ("dumpConvexity", C.c_bool), ("dumpWedgeMin", C.c_bool),
        ("redB2", C.c_bool), ("redB2f", C.c_bool), ("redBMax", C.c_bool), ("redJMax", C.c_bool),
        ("redJxB_B2Max", C.c_bool), ("redEpsilonStar", C.c_bool), ("redErrB_1ez", C.c_bool),
        ("redErrXb_XbAn", C.c_bool), ("redConvex", C.c_bool), ("redWedgeMin", C.c_bool),
        ("redU2", C.c_bool), ("redUMax", C.c_bool),
    ]

    def set_str(self, name, value):
        """Fortran character(len=30) as copyString leaves it: blank padded,
        29 characters and a NUL (src/readNamelist.f90:267-283)."""
        b = value.encode()[:29].ljust(29, b" ") + b"\0"
        C.memmove(C.addressof(self) + getattr(Params, name).offset, b, 30)

    def get_str(self, name):
        raw = bytes(C.string_at(C.addressof(self) + getattr(Params, name).offset, 30))
        return raw.split(b"\0")[0].decode().rstrip()

    @property
    def npad(self):
        return (self.nx + 2) * (self.ny + 2) * (self.nz + 2)

    @property
    def ninner(self):
        return self.nx * self.ny * self.nz


class Reduction(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "B2", "B2res", "BMax", "JMax", "JxB_B2Max", "epsilonStar", "errB_1ez", "errXb_XbAn",
        "convex", "wedgeMin", "U2", "UMax", "JxB_JBMax", "JxBMax_JBMax")]


class StepInfo(C.Structure):
    _fields_ = [("accepted", C.c_int), ("nan_seen", C.c_int), ("stopped", C.c_int), ("converged", C.c_int),
                ("dt_used", C.c_double), ("maxDelta", C.c_double), ("t", C.c_double), ("dt", C.c_double),
                ("n_attempt", C.c_int64), ("n_accept", C.c_int64)]


class RunOpts(C.Structure):
    _fields_ = [("max_attempts", C.c_int64), ("n_cycle", C.c_int), ("stop_ratio", C.c_double),
                ("batch", C.c_int), ("use_graph", C.c_int), ("dt_dump", C.c_double), ("dt_save", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("n_attempt", C.c_int64), ("n_accept", C.c_int64), ("t", C.c_double), ("dt", C.c_double),
                ("ratio", C.c_double), ("ratio_pt", C.c_double), ("energy", C.c_double),
                ("converged", C.c_int), ("stopped", C.c_int),
                ("nan_seen", C.c_int), ("seconds", C.c_double), ("event", C.c_int)]


class StepRecord(C.Structure):
    _fields_ = [("t", C.c_double), ("dt_used", C.c_double), ("maxDelta", C.c_double), ("ratio", C.c_double),
                ("ratio_pt", C.c_double), ("energy", C.c_double), ("jxb_b2", C.c_double), ("jmax", C.c_double),
                ("accepted", C.c_int32), ("nan_seen", C.c_int32)]


class WarmInfo(C.Structure):
    _fields_ = [("used", C.c_int), ("mesh_ok", C.c_int), ("settle_attempts", C.c_int64),
                ("ratio_cold", C.c_double), ("ratio_applied", C.c_double), ("ratio_warm", C.c_double),
                ("energy_cold", C.c_double), ("energy_applied", C.c_double), ("energy_warm", C.c_double),
                ("min_det", C.c_double), ("ms", C.c_double)]


WARM_FORCE = 1


_P = C.POINTER
_D = _P(C.c_double)
_CTX = C.c_void_p

_SIGS = {
    "gm_last_error": (C.c_char_p, []),
    "gm_version": (C.c_int, []),
    "gm_params_default": (C.c_int, [_P(Params)]),
    "gm_params_read": (C.c_int, [C.c_char_p, _P(Params)]),
    "gm_params_derive": (C.c_int, [_P(Params)]),
    "gm_init_state": (C.c_int, [_P(Params), _D, _D]),
    "gm_init_residual": (C.c_double, [_P(Params)]),
    "gm_braid_seeded": (C.c_int, [_P(Params), C.c_uint64, C.c_int, C.c_int, C.c_int]),
    "gm_create": (C.c_int, [_P(Params), C.c_int, _P(_CTX)]),
    "gm_destroy": (C.c_int, [_CTX]),
    "gm_upload": (C.c_int, [_CTX, _D, _D, C.c_double, C.c_double]),
    "gm_upload_velocity": (C.c_int, [_CTX, _D]),
    "gm_download": (C.c_int, [_CTX, C.c_int, _D]),
    "gm_get_time": (C.c_int, [_CTX, _D, _D]),
    "gm_set_time": (C.c_int, [_CTX, C.c_double, C.c_double]),
    "gm_synchronize": (C.c_int, [_CTX]),
    "gm_b_jacb0": (C.c_int, [_CTX]),
    "gm_set_bbound": (C.c_int, [_CTX]),
    "gm_set_peri": (C.c_int, [_CTX, C.c_int]),
    "gm_set_grid_peri": (C.c_int, [_CTX]),
    "gm_current": (C.c_int, [_CTX]),
    "gm_gradp": (C.c_int, [_CTX]),
    "gm_rk_kk": (C.c_int, [_CTX, C.c_int, C.c_double]),
    "gm_rk_xnewstar": (C.c_int, [_CTX, _D]),
    "gm_rk_xnewstar_noalloc": (C.c_int, [_CTX, _D]),
    "gm_accept": (C.c_int, [_CTX]),
    "gm_reductions": (C.c_int, [_CTX, _P(Reduction)]),
    "gm_fields": (C.c_int, [_CTX]),
    "gm_step": (C.c_int, [_CTX, _P(StepInfo)]),
    "gm_run": (C.c_int, [_CTX, _P(RunOpts), _P(RunStats)]),
    "gm_step_log": (C.c_int, [_CTX, _P(StepRecord), C.c_int]),
    "gm_timing_enable": (C.c_int, [_CTX, C.c_int]),
    "gm_timing_read": (C.c_int, [_CTX, _D, _P(C.c_int64)]),
    "gm_algorithmic_bytes_per_point_stage": (C.c_double, [C.c_int]),
    "gm_set_tile": (C.c_int, [_CTX, C.c_int, C.c_int, C.c_int]),
    "gm_copy_probe": (C.c_int, [C.c_int64, C.c_int, _D]),
    "gm_hbm_probe": (C.c_int, [C.c_int64, C.c_int, C.c_int, _D]),
    "gm_get_tile": (C.c_int, [_CTX, _P(C.c_int), _P(C.c_int), _P(C.c_int), _P(C.c_int)]),
    "gm_write_b0": (C.c_int, [C.c_char_p, _P(Params), _D, C.c_int]),
    "gm_dump_state": (C.c_int, [C.c_char_p, _P(Params), C.c_double, C.c_double, C.c_int,
                                _D, _D, _D, _D, _D, _D, _D, _D, C.c_int]),
    "gm_read_state": (C.c_int, [C.c_char_p, _P(Params), _D, _D, _D, _D, _D, C.c_int]),
    "gm_write_ts": (C.c_int, [C.c_char_p, _P(Params), _P(Reduction), C.c_double, C.c_double, C.c_int,
                               C.c_double, C.c_int, C.c_int]),
    "gm_write_device_info": (C.c_int, [C.c_char_p, C.c_int, C.c_int]),
    "gm_get_stage_tile": (C.c_int, [_CTX, C.c_int, _P(C.c_int), _P(C.c_int), _P(C.c_int), _P(C.c_int)]),
    "gm_autotune": (C.c_int, [_CTX, C.c_int, _D]),
    "gm_set_stage_tile": (C.c_int, [_CTX, C.c_int, C.c_int, C.c_int, C.c_int]),
    "gm_batch_autotune": (C.c_int, [_CTX, C.c_int, _D]),
    "gm_batch_timing_enable": (C.c_int, [_CTX, C.c_int]),
    "gm_batch_timing_read": (C.c_int, [_CTX, _D, _P(C.c_int64)]),
    "gm_batch_create": (C.c_int, [_P(Params), C.c_int, C.c_int, _P(_CTX)]),
    "gm_batch_destroy": (C.c_int, [_CTX]),
    "gm_batch_context": (C.c_int, [_CTX, C.c_int, _P(_CTX)]),
    "gm_batch_run": (C.c_int, [_CTX, _P(RunOpts), _P(RunStats)]),
    "gm_unet_num_params": (C.c_int64, []),
    "gm_unet_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _P(_CTX)]),
    "gm_unet_destroy": (C.c_int, [_CTX]),
    "gm_unet_init_params": (C.c_int, [_CTX, C.c_uint64]),
    "gm_unet_set_params": (C.c_int, [_CTX, _D, C.c_int64]),
    "gm_unet_get_params": (C.c_int, [_CTX, _D, C.c_int64]),
    "gm_unet_forward": (C.c_int, [_CTX, _D, _D]),
    "gm_unet_loss_grad": (C.c_int, [_CTX, _D, _D, _D, _D]),
    "gm_unet_add_sample": (C.c_int, [_CTX, _CTX]),
    "gm_unet_num_samples": (C.c_int, [_CTX]),
    "gm_unet_clear_samples": (C.c_int, [_CTX]),
    "gm_unet_train": (C.c_int, [_CTX, C.c_int, C.c_int, C.c_double, C.c_uint64, _D]),
    "gm_warm_start": (C.c_int, [_CTX, _CTX, C.c_int, C.c_int, _P(WarmInfo)]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

lib = _lib
EXPORTED = tuple(_SIGS)


class GlemurError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _check(code):
    if code < 0:
        raise GlemurError(code, _lib.gm_last_error().decode())
    return code


def _dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def params_default():
    p = Params()
    _check(_lib.gm_params_default(C.byref(p)))
    return p


def params_read(path, base=None):
    p = params_default() if base is None else base
    _check(_lib.gm_params_read(str(path).encode(), C.byref(p)))
    return p


def params_derive(p):
    _check(_lib.gm_params_derive(C.byref(p)))
    return p


def init_state(p):
    """B0 and xb (3P doubles each, reference AoS padded layout)."""
    B0 = np.empty(3 * p.npad)
    xb = np.empty(3 * p.npad)
    _check(_lib.gm_init_state(C.byref(p), _dptr(B0), _dptr(xb)))
    return B0, xb


def init_residual(p):
    return _lib.gm_init_residual(C.byref(p))


def braid_seeded(p, seed, nmin=6, nmax=6, random_centres=False):
    _check(_lib.gm_braid_seeded(C.byref(p), C.c_uint64(seed), nmin, nmax, int(random_centres)))
    return p


def copy_probe(nbytes, reps=10):
    """Mean ms of a plain device copy of nbytes (reads nbytes, writes nbytes)."""
    ms = C.c_double()
    _check(_lib.gm_copy_probe(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def hbm_probe(nbytes, mode, reps=10):
    """Mean ms of a device copy (mode 0: reads and writes nbytes), read
    (1: nbytes) or write (2: nbytes) with eight 16-byte words in flight per lane."""
    ms = C.c_double()
    _check(_lib.gm_hbm_probe(int(nbytes), int(mode), int(reps), C.byref(ms)))
    return ms.value


def algorithmic_bytes(precision):
    return _lib.gm_algorithmic_bytes_per_point_stage(precision)


class Context:
    """One relaxation on the current HIP device (gm_create ... gm_destroy)."""

    def __init__(self, params, precision=PREC_F64):
        self.p = params
        self.precision = precision
        self._owned = True
        self._h = _CTX()
        _check(_lib.gm_create(C.byref(params), precision, C.byref(self._h)))

    @classmethod
    def _member(cls, handle, params, precision):
        """A view of an ensemble member (owned by its Batch)."""
        m = cls.__new__(cls)
        m.p = params
        m.precision = precision
        m._owned = False
        m._h = handle
        return m

    def close(self):
        if self._h and self._owned:
            _lib.gm_destroy(self._h)
        self._h = _CTX()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state
    def upload(self, B0, xb, t=0.0, dt=None):
        B0 = np.ascontiguousarray(B0, dtype=np.float64)
        xb = np.ascontiguousarray(xb, dtype=np.float64)
        assert B0.size == 3 * self.p.npad and xb.size == 3 * self.p.npad
        _check(_lib.gm_upload(self._h, _dptr(B0), _dptr(xb), t, self.p.dt0 if dt is None else dt))

    def upload_velocity(self, uu):
        """Velocity of the inner points (inner AoS, 3N); inertia only."""
        uu = np.ascontiguousarray(uu, dtype=np.float64)
        assert uu.size == 3 * self.p.ninner
        _check(_lib.gm_upload_velocity(self._h, _dptr(uu)))

    def download(self, field):
        sizes = {FIELD_XB: 3 * self.p.npad, FIELD_B0: 3 * self.p.npad, FIELD_B: 3 * self.p.npad,
                 FIELD_DETJAC: self.p.npad, FIELD_J: 3 * self.p.ninner, FIELD_GRADP: 3 * self.p.ninner,
                 FIELD_KK: (36 if self.p.inertia else 18) * self.p.ninner, FIELD_XB_NEW: 3 * self.p.npad,
                 FIELD_XB_LO: 3 * self.p.npad, FIELD_CELLVOL: self.p.ninner, FIELD_CONVEXITY: self.p.ninner,
                 FIELD_WEDGEMIN: self.p.ninner, FIELD_UU: 3 * self.p.ninner}
        out = np.empty(sizes[field])
        _check(_lib.gm_download(self._h, field, _dptr(out)))
        return out

    def time(self):
        t, dt = C.c_double(), C.c_double()
        _check(_lib.gm_get_time(self._h, C.byref(t), C.byref(dt)))
        return t.value, dt.value

    def set_time(self, t, dt):
        _check(_lib.gm_set_time(self._h, t, dt))

    def sync(self):
        _check(_lib.gm_synchronize(self._h))

    # -- the reference's call structure, one kernel per call
    def b_jacb0(self): _check(_lib.gm_b_jacb0(self._h))
    def set_bbound(self): _check(_lib.gm_set_bbound(self._h))
    def set_peri_B(self): _check(_lib.gm_set_peri(self._h, FIELD_B))
    def set_grid_peri(self): _check(_lib.gm_set_grid_peri(self._h))
    def current(self): _check(_lib.gm_current(self._h))
    def gradp(self): _check(_lib.gm_gradp(self._h))
    def rk_kk(self, n, dt): _check(_lib.gm_rk_kk(self._h, n, dt))
    def accept(self): _check(_lib.gm_accept(self._h))

    def rk_xnewstar(self):
        md = C.c_double()
        _check(_lib.gm_rk_xnewstar(self._h, C.byref(md)))
        return md.value

    def rk_xnewstar_noalloc(self):
        """gm_rk_xnewstar_noalloc: the same results with the library's static
        scratch instead of allocated memory."""
        md = C.c_double()
        _check(_lib.gm_rk_xnewstar_noalloc(self._h, C.byref(md)))
        return md.value

    def reductions(self):
        r = Reduction()
        _check(_lib.gm_reductions(self._h, C.byref(r)))
        return r

    # -- fused path
    def fields(self):
        _check(_lib.gm_fields(self._h))

    def step(self):
        s = StepInfo()
        _check(_lib.gm_step(self._h, C.byref(s)))
        return s

    def run(self, max_attempts=0, n_cycle=0, stop_ratio=0.0, batch=16, use_graph=False):
        o = RunOpts(max_attempts, n_cycle, stop_ratio, batch, int(use_graph))
        st = RunStats()
        _check(_lib.gm_run(self._h, C.byref(o), C.byref(st)))
        return st

    def step_log(self, n):
        buf = (StepRecord * n)()
        m = _check(_lib.gm_step_log(self._h, buf, n))
        return [buf[i] for i in range(m)]

    def timing(self, on=True):
        _check(_lib.gm_timing_enable(self._h, int(on)))

    def timing_read(self):
        """(ms, count) summed per kind: stages 0..5, then the controller."""
        ms = (C.c_double * TIMING_KINDS)()
        cnt = (C.c_int64 * TIMING_KINDS)()
        _check(_lib.gm_timing_read(self._h, ms, cnt))
        return list(ms), list(cnt)

    def set_tile(self, tx, ty, cz):
        _check(_lib.gm_set_tile(self._h, tx, ty, cz))

    def stage_tile(self, stage):
        """(tx, ty, cz, workgroups) of one stage kernel (-1: fields pass)."""
        v = [C.c_int() for _ in range(4)]
        _check(_lib.gm_get_stage_tile(self._h, stage, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_stage_tile(self, stage, tx, ty, cz):
        """Tile of one stage kernel (-1: fields pass): tx x ty columns, cz planes."""
        _check(_lib.gm_set_stage_tile(self._h, stage, tx, ty, cz))

    def autotune(self, reps=3):
        """Pick each stage kernel's tile by timing candidates; returns the
        winners' mean launch times (ms) for stages 0..5."""
        ms = (C.c_double * 6)()
        _check(_lib.gm_autotune(self._h, int(reps), ms))
        return list(ms)

    def tile(self):
        v = [C.c_int() for _ in range(4)]
        _check(_lib.gm_get_tile(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)


class Batch:
    """An ensemble of relaxations advanced together (gm_batch_*, FP32 with
    compensated positions or FP64).  Members are Context views: upload,
    download and tiles per member; run() advances all of them."""

    def __init__(self, params_list, precision=PREC_F32K):
        self.params = list(params_list)
        arr = (Params * len(self.params))(*self.params)
        self._h = _CTX()
        _check(_lib.gm_batch_create(arr, len(self.params), precision, C.byref(self._h)))
        self.members = []
        for i, p in enumerate(self.params):
            h = _CTX()
            _check(_lib.gm_batch_context(self._h, i, C.byref(h)))
            self.members.append(Context._member(h, p, precision))

    def __len__(self):
        return len(self.members)

    def __getitem__(self, i):
        return self.members[i]

    def run(self, max_attempts=0, n_cycle=0, stop_ratio=0.0, batch=16, use_graph=False):
        o = RunOpts(max_attempts, n_cycle, stop_ratio, batch, int(use_graph))
        st = (RunStats * len(self.members))()
        _check(_lib.gm_batch_run(self._h, C.byref(o), st))
        return [st[i] for i in range(len(self.members))]

    def timing(self, on=True):
        _check(_lib.gm_batch_timing_enable(self._h, int(on)))

    def timing_read(self):
        """(ms, count) summed per kind: stage launches 0..5, then the controller."""
        ms = (C.c_double * TIMING_KINDS)()
        cnt = (C.c_int64 * TIMING_KINDS)()
        _check(_lib.gm_batch_timing_read(self._h, ms, cnt))
        return list(ms), list(cnt)

    def autotune(self, reps=3):
        """Pick every stage's tile policy for the whole ensemble by timing
        candidates; returns the winners' mean launch times (ms), stages 0..5."""
        ms = (C.c_double * 6)()
        _check(_lib.gm_batch_autotune(self._h, int(reps), ms))
        return list(ms)

    def close(self):
        if self._h:
            for m in self.members:
                m._h = _CTX()
            _lib.gm_batch_destroy(self._h)
            self._h = _CTX()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def write_b0(dirpath, p, B0, precision_bits=64):
    _check(_lib.gm_write_b0(str(dirpath).encode(), C.byref(p), _dptr(np.ascontiguousarray(B0)), precision_bits))


def dump_state(dirpath, p, t, dt, n, xb, B, J=None, detJac=None, precision_bits=64, cellVol=None,
               convexity=None, wedgeMin=None, uu=None):
    nul = C.cast(None, _D)
    opt = lambda a: _dptr(a) if a is not None else nul
    _check(_lib.gm_dump_state(str(dirpath).encode(), C.byref(p), t, dt, n, _dptr(xb), _dptr(B), opt(uu), opt(J),
                              opt(detJac), opt(cellVol), opt(convexity), opt(wedgeMin), precision_bits))


def read_state(dirpath, p, precision_bits=64):
    """(xb, B0, t, dt) of a restart, and the velocity (inner AoS) as a
    fifth item when p.inertia."""
    xb = np.empty(3 * p.npad)
    B0 = np.empty(3 * p.npad)
    uu = np.empty(3 * p.ninner) if p.inertia else None
    t, dt = C.c_double(), C.c_double()
    _check(_lib.gm_read_state(str(dirpath).encode(), C.byref(p), _dptr(xb), _dptr(B0),
                              _dptr(uu) if uu is not None else C.cast(None, _D), C.byref(t), C.byref(dt),
                              precision_bits))
    if p.inertia:
        return xb, B0, t.value, dt.value, uu
    return xb, B0, t.value, dt.value


def write_ts(dirpath, p, red, t, dt, it, maxDelta, echo=False, jxb_jb_col=True):
    _check(_lib.gm_write_ts(str(dirpath).encode(), C.byref(p), C.byref(red) if red is not None else None,
                            t, dt, it, maxDelta, int(echo), int(jxb_jb_col)))


class UNet:
    """Warm-start U-Net on an nx x ny x nz grid (each 4m+1); `capacity`
    training samples are held in device memory."""

    def __init__(self, nx, ny, nz, capacity=0):
        self.shape = (nz, ny, nx)
        self._h = _CTX()
        _check(_lib.gm_unet_create(nx, ny, nz, capacity, C.byref(self._h)))

    def close(self):
        if self._h:
            _lib.gm_unet_destroy(self._h)
            self._h = _CTX()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def num_params():
        return int(_lib.gm_unet_num_params())

    def init_params(self, seed):
        _check(_lib.gm_unet_init_params(self._h, C.c_uint64(seed)))

    def set_params(self, params):
        params = np.ascontiguousarray(params, dtype=np.float64)
        _check(_lib.gm_unet_set_params(self._h, _dptr(params), params.size))

    def get_params(self):
        out = np.empty(self.num_params())
        _check(_lib.gm_unet_get_params(self._h, _dptr(out), out.size))
        return out

    def forward(self, feat):
        """feat [nz][ny][nx][8] -> prediction [nz][ny][nx][3]."""
        feat = np.ascontiguousarray(feat, dtype=np.float64)
        assert feat.shape == self.shape + (8,)
        y = np.empty(self.shape + (3,))
        _check(_lib.gm_unet_forward(self._h, _dptr(feat), _dptr(y)))
        return y

    def loss_grad(self, feat, target):
        feat = np.ascontiguousarray(feat, dtype=np.float64)
        target = np.ascontiguousarray(target, dtype=np.float64)
        assert feat.shape == self.shape + (8,) and target.shape == self.shape + (3,)
        loss = C.c_double()
        grad = np.empty(self.num_params() - 1)
        _check(_lib.gm_unet_loss_grad(self._h, _dptr(feat), _dptr(target), C.byref(loss), _dptr(grad)))
        return loss.value, grad

    def add_sample(self, ctx):
        _check(_lib.gm_unet_add_sample(self._h, ctx._h))

    def num_samples(self):
        return _lib.gm_unet_num_samples(self._h)

    def clear_samples(self):
        _check(_lib.gm_unet_clear_samples(self._h))

    def train(self, epochs, batch=8, lr=1e-3, seed=0):
        loss = C.c_double()
        _check(_lib.gm_unet_train(self._h, epochs, batch, lr, C.c_uint64(seed), C.byref(loss)))
        return loss.value

    def warm_start(self, ctx, settle=0, force=False):
        info = WarmInfo()
        _check(_lib.gm_warm_start(ctx._h, self._h, int(settle), WARM_FORCE if force else 0, C.byref(info)))
        return info
