"""ctypes access to the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only the tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this package, as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = C.CDLL(LIB_PATH)


class Params(C.Structure):
    """or_params: the reference's struct Parameters with REAL = double."""
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

    @property
    def npad(self):
        return (self.nx + 2) * (self.ny + 2) * (self.nz + 2)

    @property
    def ninner(self):
        return self.nx * self.ny * self.nz


class State(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in
                ("B0", "xb", "B", "detJac", "J", "gradP", "kk", "xb_tmp", "xb_new")] + [
        ("t", C.c_double), ("dt", C.c_double),
        ("a", C.c_double * 6), ("b", C.c_double * 42), ("c", C.c_double * 6), ("cs", C.c_double * 6)] + [
        (n, C.POINTER(C.c_double)) for n in ("uu", "uu_tmp", "uu_new")]


_D = C.POINTER(C.c_double)
_PP = C.POINTER(Params)
_SIGS = {
    "or_assign_coefficients": (None, [_D, _D, _D, _D]),
    "or_init_state": (C.c_int, [_PP, _D, _D]),
    "or_init_residual": (C.c_double, [_PP]),
    "or_b_jacb0": (None, [_PP, _D, _D, _D, _D]),
    "or_update_bbound": (None, [_PP, _D, _D, _D, _D]),
    "or_set_peri": (None, [_PP, _D]),
    "or_set_grid_peri": (None, [_PP, _D]),
    "or_jstokes": (None, [_PP, _D, _D, _D]),
    "or_jclassic": (None, [_PP, _D, _D, _D, _D]),
    "or_jstokes4th": (None, [_PP, _D, _D, _D]),
    "or_gradp": (None, [_PP, _D, _D, _D]),
    "or_kk": (None, [_PP, _D, C.c_int, C.c_double, _D, _D, _D, _D, _D, _D]),
    "or_xnewstar": (C.c_double, [_PP, _D, _D, _D, _D, _D, C.POINTER(C.c_int)]),
    "or_kk_inertia": (None, [_PP, _D, C.c_int, C.c_double, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "or_xnewstar_inertia": (C.c_double, [_PP, _D, _D, _D, _D, _D, _D, _D, C.POINTER(C.c_int)]),
    "or_u2": (C.c_double, [_PP, _D, _D]),
    "or_umax": (C.c_double, [_PP, _D]),
    "or_b2": (C.c_double, [_PP, _D, _D]),
    "or_jmax": (C.c_double, [_PP, _D]),
    "or_jxb_b2_max": (C.c_double, [_PP, _D, _D]),
    "or_jxb_jb_max": (C.c_double, [_PP, _D, _D]),
    "or_jxbmax_jbmax": (C.c_double, [_PP, _D, _D]),
    "or_mesh": (None, [_PP, _D, _D, _D, _D, _D, _D]),
    "or_err_b_1ez": (C.c_double, [_PP, _D]),
    "or_err_xb_xban": (C.c_double, [_PP, _D]),
    "or_epsilon_star": (C.c_double, [_PP, _D, _D, _D]),
    "or_state_create": (C.c_int, [_PP, _D, _D, C.POINTER(State)]),
    "or_state_free": (None, [C.POINTER(State)]),
    "or_attempt": (C.c_double, [_PP, C.POINTER(State), C.POINTER(C.c_int)]),
    "or_step": (C.c_int, [_PP, C.POINTER(State), _D, _D, C.POINTER(C.c_int)]),
    "or_fields": (None, [_PP, C.POINTER(State)]),
    "or_num_threads": (C.c_int, []),
    "or_set_threads": (None, [C.c_int]),
}
for _n, (_r, _a) in _SIGS.items():
    _f = getattr(_lib, _n)
    _f.restype = _r
    _f.argtypes = _a
lib = _lib


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def from_product(pp):
    """Copy a glemur_amd.Params (same layout) into oracle Params."""
    q = Params()
    C.memmove(C.byref(q), C.byref(pp), C.sizeof(Params))
    return q


def coefficients():
    a, b, c, cs = np.zeros(6), np.zeros(42), np.zeros(6), np.zeros(6)
    _lib.or_assign_coefficients(dp(a), dp(b), dp(c), dp(cs))
    return a, b, c, cs


def init_state(p):
    B0, xb = np.zeros(3 * p.npad), np.zeros(3 * p.npad)
    if _lib.or_init_state(C.byref(p), dp(B0), dp(xb)) != 0:
        raise ValueError("unsupported bInit")
    return B0, xb


class Sim:
    """Whole-step state of the oracle (src/main.cu:259-384)."""

    def __init__(self, p, B0, xb):
        self.p = p
        self.s = State()
        self._B0 = np.ascontiguousarray(B0, dtype=np.float64)
        self._xb = np.ascontiguousarray(xb, dtype=np.float64)
        if _lib.or_state_create(C.byref(p), dp(self._B0), dp(self._xb), C.byref(self.s)) != 0:
            raise MemoryError("oracle state")

    def __del__(self):
        try:
            _lib.or_state_free(C.byref(self.s))
        except Exception:
            pass

    def arr(self, name):
        n = {"B0": 3 * self.p.npad, "xb": 3 * self.p.npad, "B": 3 * self.p.npad, "detJac": self.p.npad,
             "J": 3 * self.p.ninner, "gradP": 3 * self.p.ninner,
             "kk": (36 if self.p.inertia else 18) * self.p.ninner,
             "xb_tmp": 3 * self.p.npad, "xb_new": 3 * self.p.npad,
             "uu": 3 * self.p.ninner, "uu_tmp": 3 * self.p.ninner, "uu_new": 3 * self.p.ninner}[name]
        if name.startswith("uu") and not self.p.inertia:
            raise ValueError("the velocity exists with inertia only")
        return np.ctypeslib.as_array(getattr(self.s, name), shape=(n,))

    @property
    def t(self):
        return self.s.t

    @property
    def dt(self):
        return self.s.dt

    def set_dt(self, dt):
        self.s.dt = dt

    def fields(self):
        _lib.or_fields(C.byref(self.p), C.byref(self.s))

    def attempt(self):
        nan = C.c_int(0)
        return _lib.or_attempt(C.byref(self.p), C.byref(self.s), C.byref(nan))

    def step(self):
        dt_used, md, nan = C.c_double(), C.c_double(), C.c_int(0)
        acc = _lib.or_step(C.byref(self.p), C.byref(self.s), C.byref(dt_used), C.byref(md), C.byref(nan))
        return bool(acc), dt_used.value, md.value, bool(nan.value)

    def gradp(self):
        _lib.or_gradp(C.byref(self.p), self.s.xb, self.s.detJac, self.s.gradP)

    def b2(self):
        return _lib.or_b2(C.byref(self.p), self.s.B, self.s.detJac)

    def jmax(self):
        return _lib.or_jmax(C.byref(self.p), self.s.J)

    def u2(self):
        return _lib.or_u2(C.byref(self.p), self.s.uu, self.s.detJac)

    def umax(self):
        return _lib.or_umax(C.byref(self.p), self.s.uu)

    def jxb_b2_max(self):
        return _lib.or_jxb_b2_max(C.byref(self.p), self.s.B, self.s.J)

    def jxb_jb_max(self):
        return _lib.or_jxb_jb_max(C.byref(self.p), self.s.B, self.s.J)

    def jxbmax_jbmax(self):
        return _lib.or_jxbmax_jbmax(C.byref(self.p), self.s.B, self.s.J)


def set_threads(n):
    _lib.or_set_threads(int(n))


def num_threads():
    return _lib.or_num_threads()


def mesh(p, xb):
    """(cellVol, convexity, wedgeMin) per inner point [N] and the grid minima
    (convexity, wedgeMin)."""
    n = p.nx * p.ny * p.nz
    cv, co, wm = np.zeros(n), np.zeros(n), np.zeros(n)
    cmin, wmin = C.c_double(), C.c_double()
    _lib.or_mesh(C.byref(p), dp(np.ascontiguousarray(xb)), dp(cv), dp(co), dp(wm), C.byref(cmin), C.byref(wmin))
    return cv, co, wm, cmin.value, wmin.value


def err_b_1ez(p, B):
    return _lib.or_err_b_1ez(C.byref(p), dp(np.ascontiguousarray(B)))


def err_xb_xban(p, xb):
    return _lib.or_err_xb_xban(C.byref(p), dp(np.ascontiguousarray(xb)))


def epsilon_star(p, xb, B, J):
    return _lib.or_epsilon_star(C.byref(p), dp(np.ascontiguousarray(xb)), dp(np.ascontiguousarray(B)),
                                dp(np.ascontiguousarray(J)))


# ---------------------------------------------------------------- U-Net
_I = C.c_int
for _n, _a in {
    "ou_conv3d": [_D, _I, _I, _I, _I, _I, _D, _D, _I, _I, _D, _I],
    "ou_conv3d_bwd_data": [_D, _I, _I, _I, _I, _I, _D, _I, _D, _I],
    "ou_conv3d_bwd_weight": [_D, _I, _D, _I, _I, _I, _I, _I, _I, _D, _D],
    "ou_pool": [_D, _I, _I, _I, _I, _I, _D, _I],
    "ou_pool_bwd": [_D, _I, _I, _I, _I, _I, _D, _I],
    "ou_up": [_D, _I, _I, _I, _I, _I, _D, _I],
    "ou_up_bwd": [_D, _I, _I, _I, _I, _I, _D, _I],
    "ou_unet_features": [_D, _I, _I, _I, _I, _I, _I, _I, _D],
    "ou_unet_loss": [_D, _D, _D, _I, _I, _I, _I, _D, _D],
}.items():
    getattr(_lib, _n).argtypes = _a
    getattr(_lib, _n).restype = None
_lib.ou_unet_loss.restype = C.c_double
_lib.ou_bf16.argtypes = [C.c_double]
_lib.ou_bf16.restype = C.c_double
_lib.ou_unet_nparams.restype = C.c_long

UNET_LAYERS = [(8, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128),
               (192, 64), (64, 64), (96, 32), (32, 32), (32, 3)]


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _opt(a):
    return None if a is None else dp(a)


def bf16(a):
    """Round to bfloat16 (nearest even), as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    fin = (u & 0x7F800000) != 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = np.where(fin, r, u.astype(np.uint32)).view(np.float32)
    return out.astype(np.float64)


def conv3d(x, w, b, relu=False):
    """x [nz][ny][nx][cin], w [cout][27][cin], b [cout] or None -> [nz][ny][nx][cout]."""
    x, w = _c(x), _c(w)
    nz, ny, nx, cin = x.shape
    cout = w.shape[0]
    out = np.zeros((nz, ny, nx, cout))
    bb = None if b is None else _c(b)
    _lib.ou_conv3d(dp(x), cin, nx, ny, nz, cin, dp(w), _opt(bb), cout, int(relu), dp(out), cout)
    return out


def conv3d_bwd_data(gy, w):
    gy, w = _c(gy), _c(w)
    nz, ny, nx, cout = gy.shape
    cin = w.shape[2]
    gx = np.zeros((nz, ny, nx, cin))
    _lib.ou_conv3d_bwd_data(dp(gy), cout, nx, ny, nz, cout, dp(w), cin, dp(gx), cin)
    return gx


def conv3d_bwd_weight(x, gy):
    x, gy = _c(x), _c(gy)
    nz, ny, nx, cin = x.shape
    cout = gy.shape[3]
    gw, gb = np.zeros((cout, 27, cin)), np.zeros(cout)
    _lib.ou_conv3d_bwd_weight(dp(x), cin, dp(gy), cout, nx, ny, nz, cin, cout, dp(gw), dp(gb))
    return gw, gb


def pool(x):
    x = _c(x)
    nz, ny, nx, c = x.shape
    out = np.zeros(((nz + 1) // 2, (ny + 1) // 2, (nx + 1) // 2, c))
    _lib.ou_pool(dp(x), c, nx, ny, nz, c, dp(out), c)
    return out


def pool_bwd(g, fine_shape):
    g = _c(g)
    nz, ny, nx = fine_shape
    c = g.shape[3]
    out = np.zeros((nz, ny, nx, c))
    _lib.ou_pool_bwd(dp(g), c, nx, ny, nz, c, dp(out), c)
    return out


def up(x):
    x = _c(x)
    nz, ny, nx, c = x.shape
    out = np.zeros((2 * nz - 1, 2 * ny - 1, 2 * nx - 1, c))
    _lib.ou_up(dp(x), c, nx, ny, nz, c, dp(out), c)
    return out


def up_bwd(g):
    g = _c(g)
    fz, fy, fx, c = g.shape
    nz, ny, nx = (fz + 1) // 2, (fy + 1) // 2, (fx + 1) // 2
    out = np.zeros((nz, ny, nx, c))
    _lib.ou_up_bwd(dp(g), c, nx, ny, nz, c, dp(out), c)
    return out


def unet_nparams():
    return int(_lib.ou_unet_nparams())


def unet_features(B0pad, fine, stride, coarse):
    """B0pad: AoS padded B0 of the fine grid (fnx, fny, fnz); features of the
    coarse grid (nx, ny, nz) sampled every `stride` fine points."""
    B0pad = _c(B0pad)
    nx, ny, nz = coarse
    f = np.zeros((nz, ny, nx, 8))
    _lib.ou_unet_features(dp(B0pad), fine[0], fine[1], fine[2], stride, nx, ny, nz, dp(f))
    return f


def unet_loss(params, feat, target, round_bf16=False, grad=True):
    """(loss, y [nz][ny][nx][3], gradient or None) of the U-Net."""
    params, feat, target = _c(params), _c(feat), _c(target)
    nz, ny, nx = feat.shape[:3]
    y = np.zeros((nz, ny, nx, 3))
    g = np.zeros(params.size - 1) if grad else None
    loss = _lib.ou_unet_loss(dp(params), dp(feat), dp(target), nx, ny, nz, int(round_bf16), dp(y), _opt(g))
    return loss, y, g


def unet_init(seed, scale=1.0):
    """He-normal weights (std sqrt(2 / (27 cin))), zero biases, head std
    scaled by 0.1; packed layer by layer, then the output scale."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (cin, cout) in enumerate(UNET_LAYERS):
        std = np.sqrt(2.0 / (27 * cin)) * (0.1 if l == len(UNET_LAYERS) - 1 else 1.0)
        parts.append(rng.standard_normal(cout * 27 * cin) * std)
        parts.append(np.zeros(cout))
    parts.append(np.array([scale]))
    return np.concatenate(parts)
