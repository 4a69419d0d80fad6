"""ctypes binding of the CPU oracle (oracle/liboracle.so) with numpy helpers.

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py."""
import ctypes as C
import os
import subprocess

import numpy as np

from se_rbf_projects_amd.ctypes_api import as_ptr, c_double_p, c_int_p

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


def lib(path=None):
    """The oracle library; `path` loads another build of the same sources (the CPU-baseline
    timing's -march=native build) in place of oracle/liboracle.so, before the first use."""
    global _lib
    if _lib is None:
        if path is None and not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(path or LIB_PATH)
        _lib.orc_max_threads.restype = C.c_int
        _lib.orc_set_threads.argtypes = [C.c_int]
        _lib.orc_householder.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, c_double_p]
        _lib.orc_dgesv.restype = C.c_int
        _lib.orc_weights.restype = C.c_int
        _lib.orc_weights_lap.restype = C.c_int
        _lib.orc_weights_var.restype = C.c_int
        _lib.orc_bf16.argtypes = [C.c_float]
        _lib.orc_bf16.restype = C.c_float
        _lib.orc_mlp_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        _lib.orc_mlp_forward.restype = C.c_double
        _lib.orc_mlp_grad.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int, C.c_double, C.c_void_p]
        _lib.orc_mlp_grad.restype = C.c_double
        _lib.orc_adam.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                  C.c_double, C.c_double, C.c_double]
        _lib.orc_closure_features.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 11 + [C.c_int, C.c_int,
                                              C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        _lib.orc_vertical.restype = C.c_int
        _lib.orc_ref_node_weights.restype = C.c_int
        for name in ("orc_velocity", "orc_velocity_cart"):
            getattr(_lib, name).argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                            C.c_double, c_double_p]
        _lib.orc_rhs.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 11 + [C.c_double, C.c_void_p]
        _lib.orc_rk4.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 9 + [C.c_int, C.c_double, C.c_double,
                                                                              C.c_void_p]
        _lib.orc_initial.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_ulonglong, C.c_void_p]
        _lib.orc_norms.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_ref_get_layers.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double]
        _lib.orc_ref_get_rbffd_DMs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]
        _lib.orc_ref_eval_partSVDotGradSV.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
                                                      C.c_void_p]
        _lib.orc_ref_eval_sumpatchSVpartSV.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                                       C.c_void_p]
        _lib.orc_ref_eval_sumpartSVpartSV.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                                      C.c_void_p]
        _lib.orc_ref_set_TC2_U_t.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        _lib.orc_ref_get_stencils.argtypes = [C.c_void_p, C.c_int]
        _lib.orc_ref_init_HH.argtypes = [C.c_void_p]
        _lib.orc_ref_update_Up.argtypes = [C.c_void_p]
        _lib.orc_ref_init_TC2.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_ref_get_solution.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.orc_ref_verify.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def knn(xyz, n):
    xyz = _c(xyz, np.float64)
    N = len(xyz)
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    idx = np.empty((N, n), np.int32)
    D = np.empty((N, n), np.float64)
    lib().orc_knn_brute(C.c_int(N), as_ptr(x), as_ptr(y), as_ptr(z), C.c_int(n), as_ptr(idx, C.c_int),
                        as_ptr(D))
    return idx, D


def householder(p, fix_pole=False):
    H = np.empty(9)
    lib().orc_householder(float(p[0]), float(p[1]), float(p[2]), int(fix_pole), as_ptr(H))
    return H.reshape(3, 3)


def dgesv(A, B):
    A = _c(A, np.float64).copy()
    B = _c(B, np.float64).copy()
    m = A.shape[0]
    nrhs = 1 if B.ndim == 1 else B.shape[1]
    ipiv = np.empty(m, np.int32)
    info = lib().orc_dgesv(C.c_int(m), C.c_int(nrhs), as_ptr(A), as_ptr(ipiv, C.c_int), as_ptr(B))
    return info, B


def weights(cfg, xyz, idx):
    xyz = _c(xyz, np.float64)
    idx = _c(idx, np.int32)
    N, n = idx.shape
    assert n == cfg.n
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    Dx, Dy, L = (np.zeros((N, n)) for _ in range(3))
    st = np.zeros(N, np.int32)
    lib().orc_weights(C.byref(cfg), C.c_int(N), as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(idx, C.c_int),
                      as_ptr(Dx), as_ptr(Dy), as_ptr(L), as_ptr(st, C.c_int))
    return Dx, Dy, L, st


def weights_lap(cfg, xyz, idx):
    """Config-mode weights plus the Delta (L1) and Delta^2 (L2) weights of the Delta^k system."""
    xyz = _c(xyz, np.float64)
    idx = _c(idx, np.int32)
    N, n = idx.shape
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    Dx, Dy, L, L1, L2 = (np.zeros((N, n)) for _ in range(5))
    st = np.zeros(N, np.int32)
    lib().orc_weights_lap(C.byref(cfg), C.c_int(N), as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(idx, C.c_int),
                          as_ptr(Dx), as_ptr(Dy), as_ptr(L), as_ptr(L1), as_ptr(L2), as_ptr(st, C.c_int))
    return Dx, Dy, L, L1, L2, st


def weights_var(cfg, xyz, idx, nn, deg):
    """Config-mode weights with per-node stencil size and degree (rows of stride cfg.n)."""
    xyz = _c(xyz, np.float64)
    idx = _c(idx, np.int32)
    N, n = idx.shape
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    Dx, Dy, L = (np.zeros((N, n)) for _ in range(3))
    st = np.zeros(N, np.int32)
    nn = _c(nn, np.int32)
    deg = _c(deg, np.int32)
    lib().orc_weights_var(C.byref(cfg), C.c_int(N), as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(idx, C.c_int),
                          as_ptr(nn, C.c_int), as_ptr(deg, C.c_int), as_ptr(Dx), as_ptr(Dy), as_ptr(L),
                          as_ptr(st, C.c_int))
    return Dx, Dy, L, st


def vertical(cfg):
    nvmax = 7 if cfg.mode == 0 else cfg.nv
    start = np.zeros(cfg.Nv, np.int32)
    W = np.zeros((cfg.Nv, nvmax))
    nv = lib().orc_vertical(C.byref(cfg), as_ptr(start, C.c_int), as_ptr(W))
    assert nv == nvmax, nv
    return start, W


def velocity(cfg, p, h, t, rotated=True):
    U = np.empty(3)
    f = lib().orc_velocity if rotated else lib().orc_velocity_cart
    f(C.byref(cfg), float(p[0]), float(p[1]), float(p[2]), float(h), float(t), as_ptr(U))
    return U


def initial(kind, cfg, xyz, seed=0):
    xyz = _c(xyz, np.float64)
    N = len(xyz)
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    q = np.empty((N, cfg.Nv))
    lib().orc_initial(kind, C.byref(cfg), N, x.ctypes.data, y.ctypes.data, z.ctypes.data, seed, q.ctypes.data)
    return q


class Problem:
    """Node set + stencils + weights + vertical operator for the generic oracle."""

    def __init__(self, cfg, xyz, idx, Dx, Dy, L):
        self.cfg = cfg
        self.xyz = _c(xyz, np.float64)
        self.x, self.y, self.z = (np.ascontiguousarray(self.xyz[:, i]) for i in range(3))
        self.idx = _c(idx, np.int32)
        self.Dx, self.Dy, self.L = (_c(a, np.float64) for a in (Dx, Dy, L))
        self.vstart, self.vW = vertical(cfg)
        self.N = len(xyz)

    def _args(self):
        return [self.x.ctypes.data, self.y.ctypes.data, self.z.ctypes.data, self.idx.ctypes.data,
                self.Dx.ctypes.data, self.Dy.ctypes.data, self.L.ctypes.data, self.vstart.ctypes.data,
                self.vW.ctypes.data]

    def rhs(self, q, t, gamma=None):
        q = _c(q, np.float64)
        F = np.empty_like(q)
        g = None if gamma is None else _c(gamma, np.float64)
        lib().orc_rhs(C.byref(self.cfg), self.N, *self._args(), None if g is None else g.ctypes.data, q.ctypes.data,
                      float(t), F.ctypes.data)
        return F

    def rk4(self, q, nsteps, dt, t0=0.0):
        q = _c(q, np.float64).copy()
        lib().orc_rk4(C.byref(self.cfg), self.N, *self._args(), int(nsteps), float(dt), float(t0), q.ctypes.data)
        return q

    def norms(self, q_ref, q):
        out = np.empty(7)
        qr = _c(q_ref, np.float64)
        qa = _c(q, np.float64)
        lib().orc_norms(C.byref(self.cfg), self.N, qr.ctypes.data, qa.ctypes.data, out.ctypes.data)
        return out


# ---- learned closure ----
NPARAMS = 4929


def bf16(x):
    return float(lib().orc_bf16(float(x)))


def mlp_forward(params, norm, feat, emulate=True):
    """Network output o for each row of feat [m, 16] (float32)."""
    p = _c(params, np.float32)
    nm = _c(norm, np.float32)
    f = _c(feat, np.float32).reshape(-1, 16)
    return np.array([lib().orc_mlp_forward(as_ptr(p, C.c_float), as_ptr(nm, C.c_float), as_ptr(f[i], C.c_float),
                                           int(emulate), None, None) for i in range(len(f))])


def mlp_grad(params, norm, feat, A, R, emulate=True, s_max=16.0, cap=None):
    """Gradient and mean loss of ((min(softplus(o), s_max, cap) - 1) A - R)^2 over the samples."""
    p = _c(params, np.float32)
    nm = _c(norm, np.float32)
    f = _c(feat, np.float32)
    A = _c(A, np.float32)
    R = _c(R, np.float32)
    cp = None if cap is None else _c(cap, np.float32)
    g = np.zeros(NPARAMS)
    loss = lib().orc_mlp_grad(as_ptr(p, C.c_float), as_ptr(nm, C.c_float), len(A), as_ptr(f, C.c_float),
                              as_ptr(A, C.c_float), as_ptr(R, C.c_float), None if cp is None else as_ptr(cp, C.c_float),
                              int(emulate), float(s_max), as_ptr(g))
    return g, loss


def adam(params, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    p = np.array(params, np.float32)
    m = np.array(m, np.float64)
    v = np.array(v, np.float64)
    g = _c(g, np.float64)
    lib().orc_adam(len(p), as_ptr(p, C.c_float), as_ptr(m), as_ptr(v), as_ptr(g), int(step), lr, b1, b2, eps)
    return p, m, v


def closure_features(cfg, xyz, idx, Dx, Dy, L, L1, L2, vstart, vW, q, t, dt):
    """Features [N*ntr*Nv, 16] and Delta^k q of q [N, ntr, Nv] (or [N, Nv])."""
    xyz = _c(xyz, np.float64)
    x, y, z = (np.ascontiguousarray(xyz[:, i]) for i in range(3))
    N, n = idx.shape
    q = _c(q, np.float64)
    ntr = q.shape[1] if q.ndim == 3 else 1
    m = N * ntr * cfg.Nv
    feat = np.zeros((m, 16), np.float32)
    lap = np.zeros(m)
    arrs = [_c(a, np.float64) for a in (Dx, Dy, L, L1, L2)]
    vs = _c(vstart, np.int32)
    vw = _c(vW, np.float64)
    lib().orc_closure_features(C.byref(cfg), N, as_ptr(x), as_ptr(y), as_ptr(z), as_ptr(_c(idx, np.int32), C.c_int),
                               *[as_ptr(a) for a in arrs], as_ptr(vs, C.c_int), as_ptr(vw), vw.shape[1], ntr,
                               as_ptr(q), float(t), float(dt), as_ptr(feat, C.c_float), as_ptr(lap))
    return feat, lap
