"""Python host binding of libserbf.so (the MI355X RBF-FD tracer-advection engine).

A thin ctypes layer over the C-ABI in include/serbf.h: device memory, kernels and
the time loop all live in the library.  Loading fails loudly when the library has
not been built; there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

from .ctypes_api import F32, F64, SerbfConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libserbf.so")

OK, ERR_ARG, ERR_HIP, ERR_SINGULAR, ERR_UNSUPPORTED, ERR_STATE, ERR_NOMEM = range(7)
IC_TC2, IC_BELL, IC_HILLS, IC_SLOTTED, IC_BELLS2, IC_CONST = range(6)


class SerbfError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where}: status {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libserbf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run se_rbf_projects_amd.build.build() first")
        L = C.CDLL(LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "serbf_last_error": ([], C.c_char_p),
            "serbf_icos_count": ([i], i),
            "serbf_icos_nodes": ([i, vp, vp, vp], i),
            "serbf_jitter_nodes": ([i, vp, vp, vp, d, C.c_ulonglong], i),
            "serbf_spatial_order": ([i, vp, vp, vp, vp], i),
            "serbf_read_nodeset": ([C.c_char_p, C.POINTER(i), vp, vp, vp, vp, vp], i),
            "serbf_lonlat": ([i, vp, vp, vp, vp, vp], None),
            "serbf_initial_condition": ([i, vp, i, vp, vp, vp, C.c_ulonglong, vp], i),
            "serbf_create": ([C.POINTER(vp), i], i),
            "serbf_destroy": ([vp], None),
            "serbf_set_nodes": ([vp, i, vp, vp, vp], i),
            "serbf_set_lonlat": ([vp, vp, vp], i),
            "serbf_knn": ([vp, i], i),
            "serbf_get_stencils": ([vp, vp, vp], i),
            "serbf_set_stencils": ([vp, i, vp], i),
            "serbf_knn_noalloc": ([i, i, vp, vp, vp, i, vp, vp], i),
            "serbf_configure": ([vp, vp], i),
            "serbf_weights": ([vp], i),
            "serbf_get_weights": ([vp, vp, vp, vp, vp], i),
            "serbf_set_weights": ([vp, vp, vp, vp], i),
            "serbf_get_vertical": ([vp, vp, vp, vp], i),
            "serbf_set_state": ([vp, vp], i),
            "serbf_get_state": ([vp, i, vp], i),
            "serbf_rhs": ([vp, d, vp], i),
            "serbf_run": ([vp, i, i, d, d], i),
            "serbf_run_timed": ([vp, i, i, d, d, vp, vp], i),
            "serbf_sync_f32_from_f64": ([vp], i),
            "serbf_norms": ([vp, i, vp, vp], i),
            "serbf_set_hv_split": ([vp, i], i),
            "serbf_set_gamma": ([vp, vp], i),
            "serbf_set_kernel_version": ([vp, i], i),
            "serbf_stream_copy": ([vp, C.c_size_t, i, vp], i),
            "serbf_set_stage_unroll": ([vp, i], i),
            "serbf_tile_info": ([vp, vp, vp], i),
            "serbf_get_lap_weights": ([vp, vp, vp], i),
            "serbf_set_stencil_sizes": ([vp, vp, vp], i),
            "serbf_weights_autotune": ([vp, i, vp, vp], i),
            "serbf_weights_set_variant": ([vp, i], i),
            "serbf_weights_stats": ([vp, vp, vp, vp], i),
            "serbf_closure_create": ([C.POINTER(vp), i], i),
            "serbf_closure_destroy": ([vp], None),
            "serbf_closure_set_params": ([vp, vp, vp], i),
            "serbf_closure_get_params": ([vp, vp], i),
            "serbf_closure_set_smax": ([vp, d], i),
            "serbf_closure_set_data": ([vp, C.c_long, vp, vp, vp], i),
            "serbf_closure_set_caps": ([vp, C.c_long, vp], i),
            "serbf_closure_grad": ([vp, C.c_long, C.c_long, vp], i),
            "serbf_closure_train": ([vp, C.c_long, C.c_long, i, i, d, vp], i),
            "serbf_closure_eval": ([vp, C.c_long, C.c_long, vp], i),
            "serbf_closure_features": ([vp, d, d, vp, vp], i),
            "serbf_closure_attach": ([vp, vp, i, d], i),
            "serbf_closure_bound": ([vp, d, vp], i),
            "serbf_closure_gamma": ([vp, i, d, d, vp], i),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _check(rc, where):
    if rc != OK:
        raise SerbfError(rc, where, lib().serbf_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def icos_nodes(level):
    """Icosahedral node set of the given subdivision level, float64 [N, 3]."""
    N = lib().serbf_icos_count(level)
    if N < 0:
        raise ValueError("level out of range")
    x, y, z = (np.empty(N) for _ in range(3))
    _check(lib().serbf_icos_nodes(level, _p(x), _p(y), _p(z)), "icos_nodes")
    return np.stack([x, y, z], axis=1)


def read_nodeset(path):
    """A node-set file of the reference (nodesets/<name>.nc), read without netCDF: float64
    [N, 5] columns x, y, z, lambda, phi (nodesets.c:18-90 reads the same five datasets)."""
    bpath = os.fsencode(path)
    N = C.c_int(0)
    _check(lib().serbf_read_nodeset(bpath, C.byref(N), None, None, None, None, None), "read_nodeset")
    cols = [np.empty(N.value) for _ in range(5)]
    n = C.c_int(N.value)
    _check(lib().serbf_read_nodeset(bpath, C.byref(n), *[_p(c) for c in cols]), "read_nodeset")
    return np.stack(cols, axis=1)


def jitter_nodes(xyz, frac, seed):
    x, y, z = (np.ascontiguousarray(xyz[:, k], dtype=np.float64).copy() for k in range(3))
    _check(lib().serbf_jitter_nodes(len(x), _p(x), _p(y), _p(z), float(frac), int(seed)), "jitter_nodes")
    return np.stack([x, y, z], axis=1)


def spatial_order(xyz):
    x, y, z = (np.ascontiguousarray(xyz[:, k], dtype=np.float64) for k in range(3))
    perm = np.empty(len(x), np.int32)
    _check(lib().serbf_spatial_order(len(x), _p(x), _p(y), _p(z), _p(perm)), "spatial_order")
    return perm


def initial_condition(kind, cfg, xyz, seed=0):
    x, y, z = (np.ascontiguousarray(xyz[:, k], dtype=np.float64) for k in range(3))
    q = np.empty((len(x), cfg.Nv))
    _check(lib().serbf_initial_condition(kind, C.byref(cfg), len(x), _p(x), _p(y), _p(z), int(seed), _p(q)),
           "initial_condition")
    return q


class Closure:
    """Learned hyperviscosity closure: parameters, training set and Adam state on one GPU."""

    NPARAMS = 4929

    def __init__(self, device=0):
        self._c = C.c_void_p()
        _check(lib().serbf_closure_create(C.byref(self._c), int(device)), "closure_create")

    def close(self):
        if self._c:
            lib().serbf_closure_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, norm):
        p = np.ascontiguousarray(params, np.float32)
        nm = np.ascontiguousarray(norm, np.float32)
        if p.size != self.NPARAMS or nm.size != 32:
            raise ValueError(f"closure parameters: {self.NPARAMS} weights and 32 normalisation values expected, "
                             f"got {p.size} and {nm.size}")
        _check(lib().serbf_closure_set_params(self._c, _p(p), _p(nm)), "closure_set_params")

    def set_smax(self, s_max):
        _check(lib().serbf_closure_set_smax(self._c, float(s_max)), "closure_set_smax")

    def params(self):
        p = np.empty(self.NPARAMS, np.float32)
        _check(lib().serbf_closure_get_params(self._c, _p(p)), "closure_get_params")
        return p

    def set_data(self, feat, A, R):
        self._feat = np.ascontiguousarray(feat, np.float32)
        self._A = np.ascontiguousarray(A, np.float32)
        self._R = np.ascontiguousarray(R, np.float32)
        n = len(self._A)
        if self._feat.shape != (n, 16) or self._A.shape != (n,) or self._R.shape != (n,):
            raise ValueError(f"closure data: feat (n, 16), A (n,), R (n,) expected, got {self._feat.shape}, "
                             f"{self._A.shape}, {self._R.shape}")
        _check(lib().serbf_closure_set_data(self._c, n, _p(self._feat), _p(self._A), _p(self._R)), "closure_set_data")
        self.n = n

    def set_caps(self, cap):
        """Per-sample clamps of the training loss for the current samples (None removes them)."""
        if cap is None:
            _check(lib().serbf_closure_set_caps(self._c, 0, None), "closure_set_caps")
            return
        self._cap = np.ascontiguousarray(cap, np.float32)
        if self._cap.shape != (self.n,):
            raise ValueError(f"closure caps: one per sample ({self.n}) expected, got {self._cap.shape}")
        _check(lib().serbf_closure_set_caps(self._c, self.n, _p(self._cap)), "closure_set_caps")

    def grad(self, first, count):
        g = np.empty(self.NPARAMS + 1, np.float32)
        _check(lib().serbf_closure_grad(self._c, int(first), int(count), _p(g)), "closure_grad")
        return g[:-1], float(g[-1])

    def train(self, steps, batch, lr, first=0, count=None):
        """Adam steps over batches of the sample range [first, first + count) (default: all)."""
        count = self.n - first if count is None else count
        losses = np.empty(max(steps, 1), np.float32)
        _check(lib().serbf_closure_train(self._c, int(first), int(count), int(steps), int(batch), float(lr),
                                         _p(losses)), "closure_train")
        return losses[:steps]

    def eval(self, first, count):
        o = np.empty(count, np.float32)
        _check(lib().serbf_closure_eval(self._c, int(first), int(count), _p(o)), "closure_eval")
        return o


class Engine:
    """Device-resident RBF-FD advection engine on one GPU."""

    def __init__(self, device=0):
        self._e = C.c_void_p()
        _check(lib().serbf_create(C.byref(self._e), int(device)), "create")
        self.N = 0
        self.n = 0
        self.cfg = None

    def close(self):
        if self._e:
            lib().serbf_destroy(self._e)
            self._e = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_nodes(self, xyz, lonlat=None):
        xyz = _f64(xyz)
        if xyz.ndim != 2 or xyz.shape[1] != 3 or len(xyz) == 0:
            raise ValueError(f"set_nodes: (N, 3) coordinates expected, got {xyz.shape}")
        if lonlat is not None:
            lam, phi = (_f64(a) for a in lonlat)
            if lam.size != len(xyz) or phi.size != len(xyz):
                raise ValueError("set_nodes: lon/lat need one value per node")
        self._xyz = [np.ascontiguousarray(xyz[:, k]) for k in range(3)]
        self.N = 0
        _check(lib().serbf_set_nodes(self._e, len(xyz), *[_p(a) for a in self._xyz]), "set_nodes")
        self.N = len(xyz)
        self.n = 0
        if lonlat is not None:
            _check(lib().serbf_set_lonlat(self._e, _p(lam), _p(phi)), "set_lonlat")

    def knn(self, n):
        _check(lib().serbf_knn(self._e, int(n)), "knn")
        self.n = n

    def stencils(self):
        idx = np.empty((self.N, self.n), np.int32)
        D = np.empty((self.N, self.n))
        _check(lib().serbf_get_stencils(self._e, _p(idx), _p(D)), "get_stencils")
        return idx, D

    def set_stencils(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        if idx.ndim != 2 or idx.shape[0] != self.N:
            raise ValueError(f"set_stencils: (N = {self.N}, n) indices expected, got {idx.shape}")
        _check(lib().serbf_set_stencils(self._e, idx.shape[1], _p(idx)), "set_stencils")
        self.n = idx.shape[1]

    def configure(self, cfg):
        if not isinstance(cfg, SerbfConfig):
            raise TypeError("configure: a SerbfConfig is expected")
        self.cfg = None   # the library leaves the engine unconfigured when this fails
        _check(lib().serbf_configure(self._e, C.byref(cfg)), "configure")
        self.cfg = cfg

    @property
    def ntr(self):
        return max(1, self.cfg.ntracers)

    def weights(self, allow_singular=False):
        rc = lib().serbf_weights(self._e)
        if not (allow_singular and rc == ERR_SINGULAR):
            _check(rc, "weights")
        return rc

    def set_stencil_sizes(self, nn, deg):
        """Per-node stencil size and polynomial degree (None, None: uniform)."""
        if nn is None:
            _check(lib().serbf_set_stencil_sizes(self._e, None, None), "set_stencil_sizes")
            return
        self._nn = np.ascontiguousarray(nn, np.int32)
        self._deg = np.ascontiguousarray(deg, np.int32)
        if self._nn.size != self.N or self._deg.size != self.N:
            raise ValueError("set_stencil_sizes: one size and one degree per node")
        _check(lib().serbf_set_stencil_sizes(self._e, _p(self._nn), _p(self._deg)), "set_stencil_sizes")

    def weights_autotune(self, sample=16384):
        """Times the weight-solve schedule variants on a node sample; returns
        [(variant, ms, systems/s)] and keeps the fastest for weights()."""
        rep = np.zeros(3 * 16)
        nv = C.c_int()
        _check(lib().serbf_weights_autotune(self._e, int(sample), _p(rep), C.byref(nv)), "weights_autotune")
        return [(int(rep[3 * k]), float(rep[3 * k + 1]), float(rep[3 * k + 2])) for k in range(nv.value)]

    def weights_variant(self, v):
        """Pins the schedule variant of weights() (0-5, as weights_autotune numbers them)."""
        _check(lib().serbf_weights_set_variant(self._e, int(v)), "weights_set_variant")

    def weights_stats(self):
        """(kernel ms, systems solved, variant) of the last weights() call."""
        ms, sy, v = C.c_double(), C.c_double(), C.c_int()
        _check(lib().serbf_weights_stats(self._e, C.byref(ms), C.byref(sy), C.byref(v)), "weights_stats")
        return ms.value, sy.value, v.value

    def get_weights(self):
        Dx, Dy, L = (np.empty((self.N, self.n)) for _ in range(3))
        st = np.empty(self.N, np.int32)
        _check(lib().serbf_get_weights(self._e, _p(Dx), _p(Dy), _p(L), _p(st)), "get_weights")
        return Dx, Dy, L, st

    def set_weights(self, Dx, Dy, L=None):
        Dx, Dy = _f64(Dx), _f64(Dy)
        L = None if L is None else _f64(L)
        for w in (Dx, Dy) if L is None else (Dx, Dy, L):
            if w.size != self.N * self.n:
                raise ValueError(f"set_weights: N x n = {self.N} x {self.n} weights per operator expected, got {w.size}")
        _check(lib().serbf_set_weights(self._e, _p(Dx), _p(Dy), _p(L)), "set_weights")

    def vertical(self):
        Nv = self.cfg.Nv
        start = np.empty(Nv, np.int32)
        W = np.empty(Nv * 9)
        nv = C.c_int()
        _check(lib().serbf_get_vertical(self._e, _p(start), _p(W), C.byref(nv)), "get_vertical")
        return start, W[: Nv * nv.value].reshape(Nv, nv.value)

    def set_state(self, q):
        q = _f64(q)
        if self.cfg is None or q.size != self.N * self.ntr * self.cfg.Nv:
            raise ValueError("set_state: N x tracers x Nv values of a configured engine expected")
        _check(lib().serbf_set_state(self._e, _p(q)), "set_state")

    def state(self, precision=F64):
        q = np.empty((self.N, self.ntr, self.cfg.Nv)) if self.ntr > 1 else np.empty((self.N, self.cfg.Nv))
        _check(lib().serbf_get_state(self._e, int(precision), _p(q)), "get_state")
        return q

    def rhs(self, t):
        F = np.empty((self.N, self.ntr, self.cfg.Nv)) if self.ntr > 1 else np.empty((self.N, self.cfg.Nv))
        _check(lib().serbf_rhs(self._e, float(t), _p(F)), "rhs")
        return F

    def run(self, precision, nsteps, dt, t0=0.0):
        _check(lib().serbf_run(self._e, int(precision), int(nsteps), float(dt), float(t0)), "run")

    def run_history(self, precision, nsteps, dt, every, t0=0.0):
        """RK4 steps with a history, as the reference prototype records q_hist
        (PY/Solver.ipynb:669-773): the state at t0 and after every `every` steps.  Steps past the
        last multiple of `every` are run but not recorded.  Returns times [H] and the states
        [H, N, (tracers,) Nv]; the stepping between snapshots stays on the device.  An attached
        closure keeps its refresh cadence across the snapshots (the engine counts steps since
        the state was set), so the snapshots and the final state equal one run of nsteps."""
        if int(every) < 1 or int(nsteps) < 0:
            raise ValueError("run_history: every >= 1 and nsteps >= 0")
        every, nsteps = int(every), int(nsteps)
        times, hist = [float(t0)], [self.state(precision)]
        done = 0
        while done + every <= nsteps:
            self.run(precision, every, dt, t0 + done * dt)
            done += every
            times.append(t0 + done * dt)
            hist.append(self.state(precision))
        if done < nsteps:
            self.run(precision, nsteps - done, dt, t0 + done * dt)
        return np.array(times), np.stack(hist)

    def run_timed(self, precision, nsteps, dt, t0=0.0):
        ms, st = C.c_double(), C.c_double()
        _check(lib().serbf_run_timed(self._e, int(precision), int(nsteps), float(dt), float(t0), C.byref(ms),
                                     C.byref(st)), "run_timed")
        return ms.value, st.value

    def set_hv_split(self, split):
        _check(lib().serbf_set_hv_split(self._e, int(bool(split))), "set_hv_split")

    def set_gamma(self, gamma):
        g = None if gamma is None else _f64(gamma)
        if g is not None and (self.cfg is None or g.size != self.N * self.ntr * self.cfg.Nv):
            raise ValueError("set_gamma: one coefficient per point (N x tracers x Nv) expected")
        _check(lib().serbf_set_gamma(self._e, _p(g)), "set_gamma")

    def set_kernel_version(self, v):
        """Stage kernel: 7 (default) the 16-node tiles with FP32 at 4 tracers on the vector
        tile kernel (a wavefront per tracer, 16-byte state accesses) and every other tile shape
        on the matrix-core tile kernel; 5 the matrix-core tile kernel for FP32 and FP64, 1 or 4
        tracers of 64 levels in config mode; 2 column kernel (every shape, and the fallback of
        the tile kernels)."""
        _check(lib().serbf_set_kernel_version(self._e, int(v)), "set_kernel_version")

    def tile_info(self):
        nt, um = C.c_int(), C.c_int()
        _check(lib().serbf_tile_info(self._e, C.byref(nt), C.byref(um)), "tile_info")
        return nt.value, um.value

    def set_stage_unroll(self, u):
        _check(lib().serbf_set_stage_unroll(self._e, int(u)), "set_stage_unroll")

    def stream_copy(self, nbytes, reps=10):
        ms = C.c_double()
        _check(lib().serbf_stream_copy(self._e, int(nbytes), int(reps), C.byref(ms)), "stream_copy")
        return ms.value

    def lap_weights(self):
        """Delta and Delta^2 weights [N, n] (config mode, hv_order >= 3)."""
        L1, L2 = (np.empty((self.N, self.n)) for _ in range(2))
        _check(lib().serbf_get_lap_weights(self._e, _p(L1), _p(L2)), "get_lap_weights")
        return L1, L2

    def closure_features(self, t, dt):
        """Closure features of the FP64 state: feat [N*ntr*Nv, 16] float32, lap [N*ntr*Nv]."""
        m = self.N * self.ntr * self.cfg.Nv
        feat = np.empty((m, 16), np.float32)
        lap = np.empty(m)
        _check(lib().serbf_closure_features(self._e, float(t), float(dt), _p(feat), _p(lap)), "closure_features")
        return feat, lap

    def attach_closure(self, closure, every=8, s_max=8.0):
        self._closure = closure
        _check(lib().serbf_closure_attach(self._e, closure._c if closure is not None else None, int(every),
                                          float(s_max)), "closure_attach")

    def closure_bound(self, kappa):
        """Per-node stability bound of the attached closure's multiplier, kappa G_ref / G_p from
        the Delta^k row sums (kappa = 0 removes it); returns the bound per node (or None)."""
        if kappa == 0:
            _check(lib().serbf_closure_bound(self._e, 0.0, None), "closure_bound")
            return None
        b = np.empty(self.N, np.float32)
        _check(lib().serbf_closure_bound(self._e, float(kappa), _p(b)), "closure_bound")
        return b

    def closure_gamma(self, precision, t, dt):
        g = np.empty((self.N, self.ntr, self.cfg.Nv)) if self.ntr > 1 else np.empty((self.N, self.cfg.Nv))
        _check(lib().serbf_closure_gamma(self._e, int(precision), float(t), float(dt), _p(g)), "closure_gamma")
        return g

    def sync_f32(self):
        _check(lib().serbf_sync_f32_from_f64(self._e), "sync_f32")

    def norms(self, precision=F64, q_ref=None):
        out = np.empty(7)
        qr = None if q_ref is None else _f64(q_ref)
        if qr is not None and (self.cfg is None or qr.size != self.N * self.ntr * self.cfg.Nv):
            raise ValueError("norms: the reference state needs N x tracers x Nv values")
        _check(lib().serbf_norms(self._e, int(precision), _p(qr), _p(out)), "norms")
        return dict(zip(("L1", "L2", "Linf", "mass_ref", "mass", "min", "max"), out))
