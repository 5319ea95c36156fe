"""ctypes binding of the libarh C-ABI (include/arh.h).

This is the FFI stub a ctypes caller needs; tests, bench.py and the torch.autograd wrapper use it.
Device buffers come from arh_malloc (DeviceArray) or any HIP allocation (torch tensors' data_ptr()).
Raises RuntimeError with the library's message on any non-zero return code; importing works
without a GPU, using the library needs one.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libarh.so")

ARH_OK = 0
ARH_FP64 = 0
ARH_MIXED = 1
ARH_RHS_POINT = 0
ARH_RHS_PLANE = 1
ARH_ECANCEL = -7
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENAN", -4: "ECONV", -5: "ENOMEM", -6: "ESTATE", -7: "ECANCEL"}

c_int, c_double, c_void_p, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
P_D = ctypes.POINTER(ctypes.c_double)
P_I = ctypes.POINTER(ctypes.c_int)


class SolveOpts(ctypes.Structure):
    _fields_ = [("tol", c_double), ("maxit", c_int), ("true_residual", c_int), ("fail_on_noconv", c_int),
                ("hist_cap", c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iters", c_int), ("matvecs", c_int), ("refinements", c_int), ("restarts", c_int),
                ("max_err", c_double), ("max_true_res", c_double), ("iters_per_rhs", P_I), ("true_res", P_D),
                ("hist", P_D), ("hist_len", P_I), ("solve_ms", c_double), ("post_ms", c_double)]


class HMatInfo(ctypes.Structure):
    _fields_ = [("n", c_int), ("n_dense", c_int), ("n_lowrank", c_int), ("n_leaves", c_int),
                ("dense_entries", ctypes.c_int64), ("lowrank_entries", ctypes.c_int64), ("max_rank", c_int),
                ("sum_rank", ctypes.c_int64), ("bytes_fp64", ctypes.c_int64), ("bytes_fp32", ctypes.c_int64),
                ("bytes_total", ctypes.c_int64), ("assemble_ms", c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class KernelProf(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_int64), ("ms", c_double), ("bytes", c_double), ("flops", c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Profile(ctypes.Structure):
    _fields_ = [("rows_fp64", KernelProf), ("lowrank_fp64", KernelProf), ("rows_fp32", KernelProf),
                ("lowrank_fp32", KernelProf)]

    def as_dict(self):
        return {f: getattr(self, f).as_dict() for f, _ in self._fields_}


class ServerOpts(ctypes.Structure):
    _fields_ = [("workers", c_int), ("hbm_budget", ctypes.c_int64), ("device", c_int), ("maxit", c_int)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", c_int), ("nrhs", c_int), ("t_submit", c_double), ("t_start", c_double),
                ("t_done", c_double), ("footprint", ctypes.c_int64), ("iters", c_int), ("max_true_res", c_double),
                ("assemble_ms", c_double), ("solve_ms", c_double), ("post_ms", c_double), ("geom_ms", c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# (name, restype, argtypes) of every entry point declared in include/arh.h
SIGNATURES = [
    ("arh_ctx_create", c_int, [c_int, ctypes.POINTER(c_void_p)]),
    ("arh_ctx_destroy", None, [c_void_p]),
    ("arh_last_error", ctypes.c_char_p, [c_void_p]),
    ("arh_ctx_set_stream", c_int, [c_void_p, c_void_p]),
    ("arh_ctx_synchronize", c_int, [c_void_p]),
    ("arh_ctx_profile", c_int, [c_void_p, c_int]),
    ("arh_ctx_get_profile", c_int, [c_void_p, ctypes.POINTER(Profile)]),
    ("arh_device_trim", c_int, [c_int, ctypes.POINTER(c_size_t)]),
    ("arh_device_cached", c_int, [c_int, ctypes.POINTER(c_size_t)]),
    ("arh_malloc", c_int, [c_void_p, c_size_t, ctypes.POINTER(c_void_p)]),
    ("arh_free", c_int, [c_void_p, c_void_p]),
    ("arh_upload", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("arh_download", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("arh_add", c_int, [c_void_p, ctypes.c_int64, c_void_p, c_void_p]),
    ("arh_partition", c_int, [P_D, c_int, c_int, c_double, P_I, P_I, c_int, P_I, P_I]),
    ("arh_geom_create", c_int, [c_void_p, P_D, c_int, P_I, c_int, P_I, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("arh_geom_set_vertices", c_int, [c_void_p, P_D]),
    ("arh_geom_set_vertices_device", c_int, [c_void_p, c_void_p]),
    ("arh_geom_set_heights_device", c_int, [c_void_p, c_void_p, P_I, c_int, c_void_p]),
    ("arh_geom_destroy", None, [c_void_p]),
    ("arh_hmat_assemble", c_int, [c_void_p, c_double, c_double, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("arh_hmat_get_info", c_int, [c_void_p, ctypes.POINTER(HMatInfo)]),
    ("arh_hmat_ranks", c_int, [c_void_p, P_I]),
    ("arh_hmat_dense_block", c_int, [c_void_p, c_int, P_D]),
    ("arh_hmat_factors", c_int, [c_void_p, c_int, P_D, P_D, P_I]),
    ("arh_hmat_compact", c_int, [c_void_p]),
    ("arh_hmat_destroy", None, [c_void_p]),
    ("arh_hmat_matvec", c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    ("arh_solve", c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, ctypes.POINTER(SolveOpts),
                          ctypes.POINTER(SolveStats)]),
    ("arh_rhs", c_int, [c_void_p, c_double, c_int, c_int, P_D, c_void_p]),
    ("arh_receivers", c_int, [c_void_p, c_double, P_D, c_int, c_int, c_int, c_void_p, c_void_p]),
    ("arh_receivers_noalloc", c_int, [c_void_p, c_double, P_D, c_int, c_int, c_int, c_void_p, c_void_p]),
    ("arh_relief_create", c_int, [c_void_p, P_D, c_int, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("arh_relief_eval", c_int, [c_void_p, c_void_p, c_void_p]),
    ("arh_relief_adjoint", c_int, [c_void_p, c_void_p, c_void_p]),
    ("arh_relief_destroy", None, [c_void_p]),
    ("arh_qaca_create", c_int, [c_void_p, c_double, P_D, c_int, c_double, c_int, ctypes.POINTER(c_void_p)]),
    ("arh_qaca_rank", c_int, [c_void_p, P_I]),
    ("arh_qaca_apply", c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    ("arh_qaca_factors", c_int, [c_void_p, P_D, P_D]),
    ("arh_qaca_destroy", None, [c_void_p]),
    ("arh_diffusion", c_int, [c_void_p, c_int, c_int, c_void_p, P_D, c_void_p, c_void_p]),
    ("arh_shape_grad", c_int, [c_void_p, c_double, P_D, c_int, P_I, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int,
                               P_D, c_void_p, c_void_p, c_void_p]),
    ("arh_shape_grad_accumulate", c_int, [c_void_p, c_double, P_D, c_int, P_I, c_int, c_int, c_void_p, c_void_p,
                                          c_void_p, c_int, P_D, c_void_p]),
    ("arh_band_solve", c_int, [c_void_p, c_double, c_double, c_int, c_int, c_int, c_int, P_D, P_D, c_int,
                               ctypes.POINTER(SolveOpts), P_D, c_void_p, c_void_p, ctypes.POINTER(SolveStats),
                               ctypes.POINTER(HMatInfo)]),
    ("arh_server_create", c_int, [ctypes.POINTER(ServerOpts), ctypes.POINTER(c_void_p)]),
    ("arh_server_add_mesh", c_int, [c_void_p, P_D, c_int, P_I, c_int, c_int, c_double, P_I]),
    ("arh_server_submit", c_int, [c_void_p, c_int, c_double, c_double, c_int, c_int, c_int, P_D, P_D, c_int, c_double,
                                  ctypes.POINTER(ctypes.c_int64)]),
    ("arh_server_wait", c_int, [c_void_p, ctypes.c_int64, P_D, ctypes.POINTER(RequestInfo)]),
    ("arh_server_poll", c_int, [c_void_p, ctypes.c_int64, ctypes.POINTER(RequestInfo)]),
    ("arh_server_profile", c_int, [c_void_p, c_int]),
    ("arh_server_get_profile", c_int, [c_void_p, ctypes.POINTER(Profile)]),
    ("arh_server_cancel", c_int, [c_void_p, P_I]),
    ("arh_server_destroy", None, [c_void_p]),
]

_lib = None
_EXITING = False
_LIVE_SERVERS = weakref.WeakSet()


def _at_exit():
    """Servers still alive at interpreter exit are destroyed first: their queued requests are
    cancelled, running ones finish and the worker threads are joined while the HIP runtime is still
    up (worker threads issuing HIP work during the runtime's own teardown crash the process). Other
    handles hold no threads and are left to the process teardown."""
    global _EXITING
    for srv in list(_LIVE_SERVERS):
        srv.close()
    _EXITING = True


atexit.register(_at_exit)


def lib():
    """Load libarh.so (fails loudly when the HIP library has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libarh.so is not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(P_D)


def _ip(a):
    return a.ctypes.data_as(P_I)


def _check(rc, ctx=None):
    if rc != ARH_OK:
        msg = lib().arh_last_error(ctx)
        raise RuntimeError("libarh %s: %s" % (ERRORS.get(rc, rc), msg.decode() if msg else ""))


def device_trim(device=0):
    """Return the library's cached device blocks to HIP (arh_device_trim); bytes freed."""
    f = c_size_t()
    _check(lib().arh_device_trim(int(device), ctypes.byref(f)))
    return f.value


def device_cached(device=0):
    """Bytes of freed device blocks the library holds for reuse (arh_device_cached)."""
    b = c_size_t()
    _check(lib().arh_device_cached(int(device), ctypes.byref(b)))
    return b.value


def partition(C, leaf, factor):
    """arh_partition: returns perm, dense (nd,4), lowrank (na,4)."""
    C = np.ascontiguousarray(C, dtype=np.float64)
    N = len(C)
    perm = np.zeros(N, np.int32)
    nd, na = c_int(), c_int()
    _check(lib().arh_partition(_dp(C), N, int(leaf), float(factor), _ip(perm), None, 0, ctypes.byref(nd),
                               ctypes.byref(na)))
    blocks = np.zeros(4 * (nd.value + na.value), np.int32)
    _check(lib().arh_partition(_dp(C), N, int(leaf), float(factor), _ip(perm), _ip(blocks), nd.value + na.value,
                               ctypes.byref(nd), ctypes.byref(na)))
    b = blocks.reshape(-1, 4)
    return perm, b[:nd.value].copy(), b[nd.value:].copy()


def centroids(P, E):
    P = np.asarray(P, dtype=np.float64)
    E = np.asarray(E)
    return (P[E[:, 0]] + P[E[:, 1]] + P[E[:, 2]]) / 3


class Context:
    def __init__(self, device=0):
        h = c_void_p()
        _check(lib().arh_ctx_create(int(device), ctypes.byref(h)))
        self.h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "h", None) and not _EXITING:
            lib().arh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def synchronize(self):
        _check(lib().arh_ctx_synchronize(self.h), self.h)

    def profile(self, enable=True):
        _check(lib().arh_ctx_profile(self.h, int(enable)), self.h)

    def get_profile(self):
        p = Profile()
        _check(lib().arh_ctx_get_profile(self.h, ctypes.byref(p)), self.h)
        return p.as_dict()

    def set_stream(self, stream_ptr):
        _check(lib().arh_ctx_set_stream(self.h, c_void_p(stream_ptr) if stream_ptr else None), self.h)

    def empty(self, shape, dtype=np.complex128):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    def diffusion(self, y, LL, nrhs, want_u=False, want_vf=False):
        d = np.zeros(nrhs)
        u = self.empty((nrhs, LL), np.float64) if want_u else None
        vf = self.empty((nrhs, LL)) if want_vf else None
        _check(lib().arh_diffusion(self.h, LL, nrhs, c_void_p(y.ptr), _dp(d), c_void_p(u.ptr) if u else None,
                                   c_void_p(vf.ptr) if vf else None), self.h)
        return d, u, vf


class DeviceArray:
    """Device allocation made through arh_malloc, freed through arh_free."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = c_void_p()
        _check(lib().arh_malloc(ctx.h, max(self.nbytes, 16), ctypes.byref(p)), ctx.h)
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _check(lib().arh_upload(self.ctx.h, c_void_p(self.ptr), a.ctypes.data_as(c_void_p), self.nbytes), self.ctx.h)

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().arh_download(self.ctx.h, out.ctypes.data_as(c_void_p), c_void_p(self.ptr), self.nbytes),
               self.ctx.h)
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and getattr(self.ctx, "h", None) and not _EXITING:
            lib().arh_free(self.ctx.h, c_void_p(self.ptr))
            self.ptr = None


def _ptr(x):
    if isinstance(x, DeviceArray):
        return c_void_p(x.ptr)
    if hasattr(x, "data_ptr"):  # torch tensor on the GPU
        return c_void_p(x.data_ptr())
    return c_void_p(int(x))


class Geometry:
    """Mesh (reordered) + frozen partition on the device (arh_geom)."""

    def __init__(self, ctx, P, E_reordered, dense, lowrank):
        self.ctx = ctx
        self.P = np.ascontiguousarray(P, dtype=np.float64)
        self.E = np.ascontiguousarray(E_reordered, dtype=np.int32)
        self.dense = np.ascontiguousarray(dense, dtype=np.int32).reshape(-1, 4)
        self.lowrank = np.ascontiguousarray(lowrank, dtype=np.int32).reshape(-1, 4)
        self.N = len(self.E)
        self.V = len(self.P)
        blocks = np.ascontiguousarray(np.vstack([self.dense, self.lowrank]).reshape(-1), dtype=np.int32)
        h = c_void_p()
        _check(lib().arh_geom_create(ctx.h, _dp(self.P), self.V, _ip(self.E), self.N, _ip(blocks), len(self.dense),
                                     len(self.lowrank), ctypes.byref(h)), ctx.h)
        self.h = h

    @classmethod
    def from_mesh(cls, ctx, P, E, leaf=32, factor=1.5, all_dense=False):
        """set_mesh semantics: partition on the input order, reorder the elements.

        all_dense=True keeps the cluster tree's leaves but makes every leaf pair a dense block: the
        exact operator (the use_actual path), whose solution is smooth in the vertex positions."""
        perm, dense, lowrank = partition(centroids(P, E), leaf, factor)
        if all_dense:
            leaves = dense[dense[:, 0] == dense[:, 2]][:, :2]
            dense = np.array([[r[0], r[1], c[0], c[1]] for r in leaves for c in leaves], dtype=np.int32)
            lowrank = np.zeros((0, 4), np.int32)
        g = cls(ctx, P, np.asarray(E, dtype=np.int32)[perm], dense, lowrank)
        g.perm = perm
        return g

    def __del__(self):
        if getattr(self, "h", None) and not _EXITING:
            lib().arh_geom_destroy(self.h)
            self.h = None

    def set_vertices(self, P):
        self.P = np.ascontiguousarray(P, dtype=np.float64)
        _check(lib().arh_geom_set_vertices(self.h, _dp(self.P)), self.ctx.h)

    def assemble(self, k, aca_tol, max_rank=100, storage=ARH_FP64):
        h = c_void_p()
        _check(lib().arh_hmat_assemble(self.h, float(k), float(aca_tol), int(max_rank), int(storage),
                                       ctypes.byref(h)), self.ctx.h)
        return HMatrix(self, h)

    def rhs(self, k, kind, vecs):
        vecs = np.ascontiguousarray(np.asarray(vecs, dtype=np.float64).reshape(-1, 3))
        out = self.ctx.empty((len(vecs), self.N))
        _check(lib().arh_rhs(self.h, float(k), int(kind), len(vecs), _dp(vecs), _ptr(out)), self.ctx.h)
        return out

    def receivers(self, k, L, x, nrhs=1, transpose=False):
        L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 3)
        LL = len(L)
        out = self.ctx.empty((nrhs, self.N if transpose else LL))
        _check(lib().arh_receivers(self.h, float(k), _dp(L), LL, int(transpose), int(nrhs), _ptr(x), _ptr(out)),
               self.ctx.h)
        return out

    def receivers_noalloc(self, k, L, x, y, nrhs=1, transpose=False):
        """arh_receivers_noalloc into a caller-provided y (device): no device allocation in the library."""
        L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 3)
        _check(lib().arh_receivers_noalloc(self.h, float(k), _dp(L), len(L), int(transpose), int(nrhs), _ptr(x),
                                           _ptr(y)), self.ctx.h)

    def listener_aca(self, k, L, tol, max_rank=5000):
        """Listener ACA Q ~= U V (compute_ACA_listener, aca.cpp:84-134) built on the device."""
        return ListenerACA(self, k, L, tol, max_rank)

    def shape_grad(self, k, L, diff, nrhs, x, lam, vf, src_kind, src):
        L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 3)
        diff = np.ascontiguousarray(diff, dtype=np.int32)
        src = np.ascontiguousarray(np.asarray(src, dtype=np.float64).reshape(-1, 3))
        HH = len(diff)
        A, B, C = (self.ctx.empty(max(HH, 1), np.float64) for _ in range(3))
        _check(lib().arh_shape_grad(self.h, float(k), _dp(L), len(L), _ip(diff), HH, int(nrhs), _ptr(x), _ptr(lam),
                                    _ptr(vf), int(src_kind), _dp(src), _ptr(A), _ptr(B), _ptr(C)), self.ctx.h)
        return A.numpy()[:HH], B.numpy()[:HH], C.numpy()[:HH]

    def band_solve(self, k, kind, vecs, L, aca_tol=1e-4, max_rank=100, storage=ARH_FP64, tol=1e-6,
                   true_residual=True, want_x=False, want_y=False, maxit=0):
        vecs = np.ascontiguousarray(np.asarray(vecs, dtype=np.float64).reshape(-1, 3))
        L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 3)
        R, LL = len(vecs), len(L)
        d = np.zeros(R)
        x = self.ctx.empty((R, self.N)) if want_x else None
        y = self.ctx.empty((R, LL)) if want_y else None
        opts = SolveOpts(tol=tol, maxit=int(maxit), true_residual=int(true_residual), fail_on_noconv=0, hist_cap=0)
        it = np.zeros(R, np.int32)
        tr = np.zeros(R)
        st = SolveStats(iters_per_rhs=_ip(it), true_res=_dp(tr))
        info = HMatInfo()
        _check(lib().arh_band_solve(self.h, float(k), float(aca_tol), int(max_rank), int(storage), int(kind), R,
                                    _dp(vecs), _dp(L), LL, ctypes.byref(opts), _dp(d), _ptr(y) if y else None,
                                    _ptr(x) if x else None, ctypes.byref(st), ctypes.byref(info)), self.ctx.h)
        stats = {"iters": st.iters, "matvecs": st.matvecs, "refinements": st.refinements, "restarts": st.restarts,
                 "max_err": st.max_err, "max_true_res": st.max_true_res, "iters_per_rhs": it, "true_res": tr,
                 "solve_ms": st.solve_ms, "post_ms": st.post_ms}
        return d, x, y, stats, info.as_dict()


class Relief:
    """Heightfield -> differentiable-vertex heights and the adjoint (arh_relief_*), for fixed uvs."""

    def __init__(self, ctx, uvs, tex_h, tex_w):
        self.ctx = ctx
        uvs = np.ascontiguousarray(uvs, dtype=np.float64).reshape(-1, 2)
        self.n, self.shape = len(uvs), (int(tex_h), int(tex_w))
        h = c_void_p()
        _check(lib().arh_relief_create(ctx.h, _dp(uvs), self.n, int(tex_h), int(tex_w), ctypes.byref(h)), ctx.h)
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and not _EXITING:
            lib().arh_relief_destroy(self.h)
            self.h = None

    def eval(self, tex, heights):
        """heights (device, n) from tex (device, tex_h x tex_w float64); device pointers or arrays."""
        _check(lib().arh_relief_eval(self.h, _ptr(tex), _ptr(heights)), self.ctx.h)

    def adjoint(self, grad, tex_grad):
        """tex_grad (device, tex_h x tex_w, overwritten) from grad (device, n)."""
        _check(lib().arh_relief_adjoint(self.h, _ptr(grad), _ptr(tex_grad)), self.ctx.h)


class ListenerACA:
    """Low-rank listener matrix (arh_qaca): Q ~= U V, Q_lj = M(L_l, element j)."""

    def __init__(self, geom, k, L, tol, max_rank=5000):
        self.geom = geom
        L = np.ascontiguousarray(L, dtype=np.float64).reshape(-1, 3)
        self.LL = len(L)
        h = c_void_p()
        _check(lib().arh_qaca_create(geom.h, float(k), _dp(L), self.LL, float(tol), int(max_rank), ctypes.byref(h)),
               geom.ctx.h)
        self.h = h
        r = c_int()
        _check(lib().arh_qaca_rank(self.h, ctypes.byref(r)))
        self.rank = r.value

    def __del__(self):
        if getattr(self, "h", None) and not _EXITING:
            lib().arh_qaca_destroy(self.h)
            self.h = None

    def apply(self, x, nrhs=1, transpose=False):
        """y = U (V x) (x nrhs x N, y nrhs x LL); transpose: y = V^T (U^T x) (x nrhs x LL)."""
        out = self.geom.ctx.empty((nrhs, self.geom.N if transpose else self.LL))
        _check(lib().arh_qaca_apply(self.h, int(transpose), int(nrhs), _ptr(x), _ptr(out)), self.geom.ctx.h)
        return out

    def factors(self):
        """U (LL x rank), V (rank x N)."""
        r, LL, N = self.rank, self.LL, self.geom.N
        U = np.zeros(max(1, LL * r), np.complex128)
        V = np.zeros(max(1, N * r), np.complex128)
        _check(lib().arh_qaca_factors(self.h, _dp(U.view(np.float64)), _dp(V.view(np.float64))), self.geom.ctx.h)
        return U[:LL * r].reshape(r, LL).T, V[:N * r].reshape(r, N)


class HMatrix:
    """Assembled H-matrix (arh_hmat)."""

    def __init__(self, geom, h):
        self.geom = geom
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and not _EXITING:
            lib().arh_hmat_destroy(self.h)
            self.h = None

    def compact(self):
        """Free the derived product copies (FP32 copy, MFMA layouts); the FP64 operator stays."""
        _check(lib().arh_hmat_compact(self.h), self.geom.ctx.h)

    def info(self):
        i = HMatInfo()
        _check(lib().arh_hmat_get_info(self.h, ctypes.byref(i)))
        return i.as_dict()

    def ranks(self):
        r = np.zeros(len(self.geom.lowrank), np.int32)
        _check(lib().arh_hmat_ranks(self.h, _ip(r)))
        return r

    def dense_block(self, b):
        rs, re, cs, ce = self.geom.dense[b]
        out = np.zeros((ce - cs) * (re - rs), np.complex128)
        _check(lib().arh_hmat_dense_block(self.h, int(b), _dp(out.view(np.float64))), self.geom.ctx.h)
        return out.reshape(ce - cs, re - rs).T

    def factors(self, a):
        """U (m x r), V (r x n) of low-rank block a (stored crosses)."""
        rs, re, cs, ce = self.geom.lowrank[a]
        m, n = re - rs, ce - cs
        r = c_int()
        _check(lib().arh_hmat_factors(self.h, int(a), None, None, ctypes.byref(r)), self.geom.ctx.h)
        r = r.value
        U = np.zeros(max(1, m * r), np.complex128)
        V = np.zeros(max(1, n * r), np.complex128)
        _check(lib().arh_hmat_factors(self.h, int(a), _dp(U.view(np.float64)), _dp(V.view(np.float64)),
                                      ctypes.byref(c_int())), self.geom.ctx.h)
        return U[:m * r].reshape(r, m).T, V[:n * r].reshape(r, n)

    def matvec(self, x, y, nrhs=1, transpose=False, precision=ARH_FP64):
        _check(lib().arh_hmat_matvec(self.h, int(transpose), int(precision), int(nrhs), _ptr(x), _ptr(y)),
               self.geom.ctx.h)

    def solve(self, b, x, nrhs=1, transpose=False, tol=1e-6, maxit=0, true_residual=False, hist_cap=0,
              fail_on_noconv=False):
        opts = SolveOpts(tol=tol, maxit=maxit, true_residual=int(true_residual), fail_on_noconv=int(fail_on_noconv),
                         hist_cap=hist_cap)
        it = np.zeros(nrhs, np.int32)
        tr = np.zeros(nrhs)
        hist = np.zeros(max(1, nrhs * hist_cap))
        hl = np.zeros(nrhs, np.int32)
        st = SolveStats(iters_per_rhs=_ip(it), true_res=_dp(tr), hist=_dp(hist) if hist_cap else None,
                        hist_len=_ip(hl) if hist_cap else None)
        _check(lib().arh_solve(self.h, int(transpose), int(nrhs), _ptr(b), _ptr(x), ctypes.byref(opts),
                               ctypes.byref(st)), self.geom.ctx.h)
        out = {"iters": st.iters, "matvecs": st.matvecs, "refinements": st.refinements, "restarts": st.restarts,
               "max_err": st.max_err, "max_true_res": st.max_true_res, "iters_per_rhs": it, "true_res": tr,
               "solve_ms": st.solve_ms}
        if hist_cap:
            out["hist"] = [hist[j * hist_cap:j * hist_cap + hl[j]].copy() for j in range(nrhs)]
        return out

    # numpy conveniences
    def matvec_np(self, x, transpose=False, precision=ARH_FP64):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        nrhs = 1 if x.ndim == 1 else x.shape[0]
        xd = self.geom.ctx.to_device(x)
        yd = self.geom.ctx.empty(x.shape)
        self.matvec(xd, yd, nrhs, transpose, precision)
        return yd.numpy()

    def solve_np(self, b, transpose=False, **kw):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        bd = self.geom.ctx.to_device(b)
        xd = self.geom.ctx.empty(b.shape)
        st = self.solve(bd, xd, nrhs, transpose, **kw)
        return xd.numpy(), st


class Server:
    """Continuously batched band solves (arh_server_*): submit returns at once, wait blocks.

    close() (also run for every live server at interpreter exit) cancels what is still queued, lets
    running requests finish and joins the worker threads."""

    def __init__(self, workers=4, hbm_budget=0, device=0, maxit=0):
        h = c_void_p()
        opts = ServerOpts(workers=int(workers), hbm_budget=int(hbm_budget), device=int(device), maxit=int(maxit))
        _check(lib().arh_server_create(ctypes.byref(opts), ctypes.byref(h)))
        self.h = h
        self._nrhs = {}
        _LIVE_SERVERS.add(self)

    def add_mesh(self, P, E, leaf=32, factor=1.5):
        P = np.ascontiguousarray(P, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.int32)
        mid = c_int()
        _check(lib().arh_server_add_mesh(self.h, _dp(P), len(P), _ip(E), len(E), int(leaf), float(factor),
                                         ctypes.byref(mid)))
        return mid.value

    def submit(self, mesh, k, vecs, L, kind=ARH_RHS_PLANE, storage=ARH_MIXED, aca_tol=1e-4, tol=1e-6):
        vecs = np.ascontiguousarray(np.asarray(vecs, dtype=np.float64).reshape(-1, 3))
        L = np.ascontiguousarray(np.asarray(L, dtype=np.float64).reshape(-1, 3))
        rid = ctypes.c_int64()
        _check(lib().arh_server_submit(self.h, int(mesh), float(k), float(aca_tol), int(storage), int(kind), len(vecs),
                                       _dp(vecs), _dp(L), len(L), float(tol), ctypes.byref(rid)))
        self._nrhs[rid.value] = len(vecs)
        return rid.value

    def wait(self, rid):
        d = np.zeros(self._nrhs[rid])
        info = RequestInfo()
        _check(lib().arh_server_wait(self.h, ctypes.c_int64(rid), _dp(d), ctypes.byref(info)))
        return d, info.as_dict()

    def poll(self, rid):
        info = RequestInfo()
        _check(lib().arh_server_poll(self.h, ctypes.c_int64(rid), ctypes.byref(info)))
        return info.as_dict()

    def cancel(self):
        """Drop the queued requests (they end with status ARH_ECANCEL); returns how many."""
        n = c_int()
        _check(lib().arh_server_cancel(self.h, ctypes.byref(n)))
        return n.value

    def profile(self, enable=True):
        _check(lib().arh_server_profile(self.h, int(enable)))

    def get_profile(self):
        p = Profile()
        _check(lib().arh_server_get_profile(self.h, ctypes.byref(p)))
        return p.as_dict()

    def close(self):
        h, self.h = getattr(self, "h", None), None
        if h:
            _LIVE_SERVERS.discard(self)
            lib().arh_server_destroy(h)

    def __del__(self):
        if not _EXITING:
            self.close()
