"""ctypes wrapper of the CPU oracle (liboracle.so) plus the analytic rigid-sphere series.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker; the product never imports it.

Reference citations are into mickey1356/acoustic_reliefs.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
P_D = ctypes.POINTER(ctypes.c_double)
P_I = ctypes.POINTER(ctypes.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build() or make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_ctx_create.restype = c_void_p
        L.orc_ctx_create.argtypes = [P_D, c_int, P_I, c_int, P_I, c_int, P_I, c_int]
        L.orc_ctx_destroy.argtypes = [c_void_p]
        for name in ["orc_set_vertices", "orc_assemble", "orc_ranks", "orc_factors", "orc_dense_block", "orc_matvec",
                     "orc_bicgstab", "orc_dense_matrix", "orc_listener_matrix", "orc_aca_listener", "orc_rhs",
                     "orc_shape_grad", "orc_aca_dense", "orc_bicgstab_dense"]:
            getattr(L, name).restype = c_int
        L.orc_set_vertices.argtypes = [c_void_p, P_D]
        L.orc_assemble.argtypes = [c_void_p, c_double, c_double, c_int]
        L.orc_ranks.argtypes = [c_void_p, P_I]
        L.orc_factors.argtypes = [c_void_p, c_int, P_D, P_D]
        L.orc_dense_block.argtypes = [c_void_p, c_int, P_D]
        L.orc_matvec.argtypes = [c_void_p, c_int, P_D, P_D]
        L.orc_bicgstab.argtypes = [c_void_p, c_int, P_D, P_D, c_double, c_int, P_D, P_I, P_D]
        L.orc_dense_matrix.argtypes = [c_void_p, c_double, P_D]
        L.orc_listener_matrix.argtypes = [c_void_p, c_double, P_D, c_int, P_D]
        L.orc_aca_listener.argtypes = [c_void_p, c_double, P_D, c_int, c_double, c_int, P_D, P_D, P_I]
        L.orc_rhs.argtypes = [c_void_p, c_double, c_int, P_D, P_D]
        L.orc_shape_grad.argtypes = [c_void_p, c_double, P_D, c_int, P_I, c_int, P_D, P_D, P_D, c_int, P_D, P_D, P_D,
                                     P_D]
        L.orc_intermediates.argtypes = [P_D, c_int, P_I, c_int, P_D, P_D]
        L.orc_entry.argtypes = [c_double, P_D, P_D, P_D]
        L.orc_dentry.argtypes = [c_double, P_D, P_D, c_int, P_D]
        L.orc_partition.argtypes = [P_D, c_int, c_int, c_double, P_I, P_I, P_I, P_I, P_I, c_int, P_I, P_I]
        L.orc_tree_check.argtypes = [P_D, c_int, c_int]
        L.orc_aca_dense.argtypes = [c_int, c_int, P_D, c_double, c_int, c_int, P_D, P_D, P_I]
        L.orc_bicgstab_dense.argtypes = [c_int, P_D, P_D, P_D, c_double, c_int, P_D, P_I, P_D]
        L.orc_dense_solve.argtypes = [c_int, P_D, P_D, P_D]
        L.orc_diffusion.restype = c_double
        L.orc_diffusion.argtypes = [c_int, P_D, P_D]
        L.orc_num_threads.restype = c_int
        L.orc_set_threads.argtypes = [c_int]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(P_D)


def _i(a):
    return a.ctypes.data_as(P_I)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def set_threads(n):
    lib().orc_set_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


# ----------------------------------------------------------------- geometry & quadrature
def intermediates(P, E):
    """Centroids and unnormalised normals (helpers.cpp:16-32)."""
    P, E = _f64(P), _i32(E)
    C = np.zeros((len(E), 3))
    Nn = np.zeros((len(E), 3))
    _check(lib().orc_intermediates(_d(P), len(P), _i(E), len(E), _d(C), _d(Nn)))
    return C, Nn


def entry(k, p, tri):
    """One matrix entry M(p, triangle) (integrate.cpp:5-14). tri: 3x3 rows a, b, c."""
    out = np.zeros(2)
    _check(lib().orc_entry(float(k), _d(_f64(p)), _d(_f64(tri).reshape(-1)), _d(out)))
    return complex(out[0], out[1])


def dentry(k, ptri, tri, which):
    """Derivative of M(centroid(ptri), tri) w.r.t. a vertex y coordinate.

    which = 0: any vertex of the collocation triangle (restates dMij_dpy, derivatives.cpp:5-35);
    which = 1, 2, 3: vertex a, b, c of the integration triangle (dMij_dq{1,2,3}y, :37-131).
    """
    out = np.zeros(2)
    _check(lib().orc_dentry(float(k), _d(_f64(ptri).reshape(-1)), _d(_f64(tri).reshape(-1)), int(which), _d(out)))
    return complex(out[0], out[1])


def partition(C, leaf, factor):
    """Cluster tree + direct/approx partition (clustertree.cpp:49-137).

    Returns perm (new position -> original element), dir (nd,4), app (na,4) as (rs,re,cs,ce),
    leaves (nl,2).
    """
    C = _f64(C)
    N = len(C)
    cap = max(16, 8 * N)
    perm = np.zeros(N, np.int32)
    dirb = np.zeros(4 * cap, np.int32)
    appb = np.zeros(4 * cap, np.int32)
    leaves = np.zeros(2 * N, np.int32)
    nd, na, nl = c_int(), c_int(), c_int()
    _check(lib().orc_partition(_d(C), N, int(leaf), float(factor), _i(perm), _i(dirb), ctypes.byref(nd), _i(appb),
                               ctypes.byref(na), cap, _i(leaves), ctypes.byref(nl)))
    return (perm, dirb[:4 * nd.value].reshape(-1, 4).copy(), appb[:4 * na.value].reshape(-1, 4).copy(),
            leaves[:2 * nl.value].reshape(-1, 2).copy())


def tree_check(C, leaf):
    return bool(lib().orc_tree_check(_d(_f64(C)), len(C), int(leaf)))


def dense_solve(A, b):
    """Dense solve A x = b (A complex N x N). Stands in for colPivHouseholderQr (adjoint.cpp:74)."""
    A = np.asfortranarray(A, dtype=np.complex128)
    N = A.shape[0]
    flat = np.ascontiguousarray(A.reshape(-1, order="F"))
    b = _c128(b)
    x = np.zeros(N, np.complex128)
    _check(lib().orc_dense_solve(N, _d(flat.view(np.float64)), _d(b.view(np.float64)), _d(x.view(np.float64))))
    return x


def aca_dense(A, eps, max_iter=100, guard=True):
    """Partially pivoted ACA (aca.cpp:38-82; guard=False: the listener variant's unguarded pivot
    division, aca.cpp:104) of an explicit m x n matrix. Returns U (m x r), V (r x n)."""
    A = np.asarray(A, dtype=np.complex128)
    m, n = A.shape
    lim = max(0, min(max_iter, m - 1, n - 1))
    flat = np.ascontiguousarray(A.reshape(-1, order="F"))
    U = np.zeros(max(1, m * lim), np.complex128)
    V = np.zeros(max(1, lim * n), np.complex128)
    r = c_int()
    _check(lib().orc_aca_dense(m, n, _d(flat.view(np.float64)), float(eps), int(max_iter), int(guard),
                               _d(U.view(np.float64)), _d(V.view(np.float64)), ctypes.byref(r)))
    r = r.value
    return U[:m * r].reshape(r, m).T, V[:r * n].reshape(r, n)


def bicgstab_dense(A, b, tol, maxit=-1):
    """The oracle's BiCGSTAB (Eigen 3.4 semantics) on an explicit matrix: x, iterations, error, history."""
    A = np.asarray(A, dtype=np.complex128)
    N = A.shape[0]
    flat = np.ascontiguousarray(A.reshape(-1, order="F"))
    b = _c128(b)
    x = np.zeros(N, np.complex128)
    mi = 2 * N if maxit < 0 else maxit
    hist = np.zeros(mi + 2)
    it = c_int()
    err = c_double()
    _check(lib().orc_bicgstab_dense(N, _d(flat.view(np.float64)), _d(b.view(np.float64)), _d(x.view(np.float64)),
                                    float(tol), int(maxit), _d(hist), ctypes.byref(it), ctypes.byref(err)))
    return x, it.value, err.value, hist[:it.value + 1].copy()


def diffusion(yt):
    """Diffusion coefficient d and u = dd/dyt (adjoint.cpp:388-397)."""
    yt = _f64(yt)
    u = np.zeros_like(yt)
    d = lib().orc_diffusion(len(yt), _d(yt), _d(u))
    return d, u


# ----------------------------------------------------------------- reference protocol helpers
def listener_points(radius, ds):
    """Hemisphere receivers (helpers.cpp:34-65): theta from +y, (r s c, r cos, -r s sin)."""
    elev = [0.0]
    azim = [0.0]
    e = ds
    while e < 91:
        a = 0.0
        while a < 360:
            elev.append(e)
            azim.append(a)
            a += ds
        e += ds
    th = np.array(elev) * np.pi / 180.0
    ph = np.array(azim) * np.pi / 180.0
    return np.column_stack([radius * np.sin(th) * np.cos(ph), radius * np.cos(th), -radius * np.sin(th) * np.sin(ph)])


def frequencies(band, n):
    """1/3-octave log-spaced frequencies of a band (helpers.cpp:67-83)."""
    if n == 1:
        return [float(band)]
    bw = 2.0 ** (1.0 / 6.0)
    lo, hi = np.log(band / bw), np.log(band * bw)
    dl = (hi - lo) / (n - 1.0)
    return [float(np.exp(lo + i * dl)) for i in range(n)]


def wavenumber(f, c0=344.0):
    """k = 2 pi f / c (helpers.h:8-10, constants.h:13)."""
    return 2 * np.pi * f / c0


# ----------------------------------------------------------------- context (mesh + partition)
class Ctx:
    """Mesh (reordered) + frozen partition + optional assembled H-matrix (hmatrix.cpp)."""

    def __init__(self, P, E, dirb, appb):
        self.P = _f64(P)
        self.E = _i32(E)
        self.N = len(self.E)
        self.dir = _i32(dirb).reshape(-1, 4)
        self.app = _i32(appb).reshape(-1, 4)
        self._h = lib().orc_ctx_create(_d(self.P), len(self.P), _i(self.E), self.N, _i(self.dir), len(self.dir),
                                       _i(self.app), len(self.app))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_ctx_destroy(self._h)
            self._h = None

    def set_vertices(self, P):
        self.P = _f64(P)
        _check(lib().orc_set_vertices(self._h, _d(self.P)))

    def assemble(self, k, aca_tol, max_rank=100):
        _check(lib().orc_assemble(self._h, float(k), float(aca_tol), int(max_rank)))

    def ranks(self):
        r = np.zeros(len(self.app), np.int32)
        _check(lib().orc_ranks(self._h, _i(r)))
        return r

    def factors(self, a):
        rs, re, cs, ce = self.app[a]
        r = int(self.ranks()[a])
        U = np.zeros(max(1, (re - rs) * r), np.complex128)
        V = np.zeros(max(1, (ce - cs) * r), np.complex128)
        _check(lib().orc_factors(self._h, int(a), _d(U.view(np.float64)), _d(V.view(np.float64))))
        return U[:(re - rs) * r].reshape(r, re - rs).T, V[:(ce - cs) * r].reshape(r, ce - cs)

    def dense_block(self, b):
        rs, re, cs, ce = self.dir[b]
        out = np.zeros((re - rs) * (ce - cs), np.complex128)
        _check(lib().orc_dense_block(self._h, int(b), _d(out.view(np.float64))))
        return out.reshape(ce - cs, re - rs).T

    def matvec(self, x, transpose=False):
        x = _c128(x)
        y = np.zeros(self.N, np.complex128)
        _check(lib().orc_matvec(self._h, int(transpose), _d(x.view(np.float64)), _d(y.view(np.float64))))
        return y

    def bicgstab(self, b, tol, maxit=-1, transpose=False):
        b = _c128(b)
        x = np.zeros(self.N, np.complex128)
        mi = 2 * self.N if maxit < 0 else maxit
        hist = np.zeros(mi + 2)
        it = c_int()
        err = c_double()
        _check(lib().orc_bicgstab(self._h, int(transpose), _d(b.view(np.float64)), _d(x.view(np.float64)),
                                  float(tol), int(maxit), _d(hist), ctypes.byref(it), ctypes.byref(err)))
        return x, it.value, err.value, hist[:it.value + 1].copy()

    def dense_matrix(self, k):
        """M - I/2, dense (integrate.cpp:32-44)."""
        out = np.zeros(self.N * self.N, np.complex128)
        _check(lib().orc_dense_matrix(self._h, float(k), _d(out.view(np.float64))))
        return out.reshape(self.N, self.N).T

    def listener_matrix(self, k, L):
        L = _f64(L)
        out = np.zeros(len(L) * self.N, np.complex128)
        _check(lib().orc_listener_matrix(self._h, float(k), _d(L), len(L), _d(out.view(np.float64))))
        return out.reshape(self.N, len(L)).T

    def aca_listener(self, k, L, eps, max_rank=5000):
        L = _f64(L)
        LL = len(L)
        cap = min(max_rank, LL - 1, self.N - 1)
        U = np.zeros(max(1, LL * cap), np.complex128)
        V = np.zeros(max(1, cap * self.N), np.complex128)
        r = c_int()
        _check(lib().orc_aca_listener(self._h, float(k), _d(L), LL, float(eps), int(max_rank), _d(U.view(np.float64)),
                                      _d(V.view(np.float64)), ctypes.byref(r)))
        r = r.value
        return U[:LL * r].reshape(r, LL).T, V[:r * self.N].reshape(r, self.N)

    def rhs_point(self, k, src):
        out = np.zeros(self.N, np.complex128)
        _check(lib().orc_rhs(self._h, float(k), 0, _d(_f64(src)), _d(out.view(np.float64))))
        return out

    def rhs_plane(self, k, direction):
        out = np.zeros(self.N, np.complex128)
        _check(lib().orc_rhs(self._h, float(k), 1, _d(_f64(direction)), _d(out.view(np.float64))))
        return out

    def shape_grad(self, k, L, diff, x, lam, vf, src_kind, src):
        L = _f64(L)
        diff = _i32(diff)
        HH = len(diff)
        A, B, C = np.zeros(HH), np.zeros(HH), np.zeros(HH)
        _check(lib().orc_shape_grad(self._h, float(k), _d(L), len(L), _i(diff), HH, _d(_c128(x).view(np.float64)),
                                    _d(_c128(lam).view(np.float64)), _d(_c128(vf).view(np.float64)), int(src_kind),
                                    _d(_f64(src)), _d(A), _d(B), _d(C)))
        return A, B, C


def setup(P, E, leaf, factor):
    """set_mesh (acoustics3d.cpp:36-67): partition on the original order, reorder Es."""
    C, _ = intermediates(P, E)
    perm, dirb, appb, leaves = partition(C, leaf, factor)
    Er = _i32(E)[perm]
    return Er, dirb, appb, leaves


# ----------------------------------------------------------------- forward / adjoint drivers
def band_forward(ctx, k, aca_tol, solver_tol, L, q_tol, src=None, directions=None, use_actual=False, max_rank=100,
                 q_max_rank=5000):
    """Forward solves of one frequency (adjoint.cpp:60-121) for a point source or plane waves.

    Returns x (nrhs, N), y (nrhs, LL), solve info list.
    """
    if use_actual:
        A = ctx.dense_matrix(k)
        Q = ctx.listener_matrix(k, L)
    else:
        ctx.assemble(k, aca_tol, max_rank)
        Uq, Vq = ctx.aca_listener(k, L, q_tol, q_max_rank)
    rhss = []
    if directions is None:
        rhss.append(ctx.rhs_point(k, src))
    else:
        for d in directions:
            rhss.append(ctx.rhs_plane(k, d))
    xs, ys, info = [], [], []
    for b in rhss:
        if use_actual:
            x = dense_solve(A, -b)
            y = Q @ x
            info.append(None)
        else:
            x, it, err, hist = ctx.bicgstab(-b, solver_tol)
            y = Uq @ (Vq @ x)
            info.append((it, err, hist))
        xs.append(x)
        ys.append(y)
    return np.array(xs), np.array(ys), info


# ----------------------------------------------------------------- analytic rigid sphere
def _sph(n, x, derivative=False):
    from scipy.special import spherical_jn, spherical_yn
    return spherical_jn(n, x, derivative), spherical_yn(n, x, derivative)


def sphere_surface_pressure(ka, cos_theta, nterms=None):
    """Total surface pressure on a sound-hard sphere, unit plane wave, e^{+i w t} convention.

    p(a, t) = conj( i/(ka)^2 sum_n i^n (2n+1) P_n(cos t) / h_n'(ka) ), t measured from the
    propagation direction (Wronskian j_n h_n' - j_n' h_n = i/x^2). Textbook series, own code.
    """
    from scipy.special import eval_legendre
    if nterms is None:
        nterms = int(ka + 40)
    ct = np.asarray(cos_theta, dtype=np.float64)
    val = np.zeros(ct.shape, np.complex128)
    for n in range(nterms):
        dj, dy = _sph(n, ka, True)
        val += (1j ** n) * (2 * n + 1) * eval_legendre(n, ct) / (dj + 1j * dy)
    return np.conj(1j / ka ** 2 * val)


def sphere_scattered(k, a, r, cos_theta, nterms=None):
    """Scattered field of a sound-hard sphere (restates misc/analytic_sphere.py:13-26)."""
    from scipy.special import eval_legendre
    if nterms is None:
        nterms = int(k * a + 40)
    ct = np.asarray(cos_theta, dtype=np.float64)
    val = np.zeros(ct.shape, np.complex128)
    for n in range(nterms):
        dja, dya = _sph(n, k * a, True)
        jr, yr = _sph(n, k * r)
        val -= (1j ** n) * (2 * n + 1) * (dja / (dja + 1j * dya)) * eval_legendre(n, ct) * (jr + 1j * yr)
    return np.conj(val)


# ----------------------------------------------------------------- DiffBEM-equivalent band pipeline
def band_value_grad(oc, ks, L, diff=None, src=(0.0, 100.0, 0.0), aca_tol=1e-3, solver_tol=1e-3, want_grad=False,
                    use_actual=False, src_kind=0, directions=None, q_aca_tol=None):
    """bem3d::gradient for one band (adjoint.cpp:24-242).

    ks: wavenumbers of the band's frequencies. src_kind 0: point source src; 1: plane wave
    src (a direction). The listener matrix is the listener ACA U V at q_aca_tol (adjoint.cpp:101,
    177) when q_aca_tol is given and use_actual is off, else the exact matrix (adjoint.cpp:75).
    Returns d, or (d, dd/dh) when want_grad.
    """
    states = []
    yt = np.zeros(len(L))
    for k in ks:
        b = oc.rhs_point(k, src) if src_kind == 0 else oc.rhs_plane(k, src)
        if use_actual:
            A = oc.dense_matrix(k)
            x = dense_solve(A, -b)
        else:
            A = None
            oc.assemble(k, aca_tol, 100)
            x, _, _, _ = oc.bicgstab(-b, solver_tol)
        if q_aca_tol is not None and not use_actual:
            U, V = oc.aca_listener(k, L, q_aca_tol)
            Q = U @ V
        else:
            Q = oc.listener_matrix(k, L)
        y = Q @ x
        yt += np.abs(y)
        states.append((k, A, Q, x, y))
    d, u = diffusion(yt)
    if not want_grad:
        return d
    g = np.zeros(len(diff))
    for k, A, Q, x, y in states:
        vf = u * np.conj(y) / np.abs(y)
        adj = Q.T @ vf
        if use_actual:
            lam = dense_solve(A.T.copy(), adj)
        else:
            oc.assemble(k, aca_tol, 100)
            lam, _, _, _ = oc.bicgstab(adj, solver_tol, transpose=True)
        gA, gB, gC = oc.shape_grad(k, L, diff, x, lam, vf, src_kind, src)
        g += gA - gB + gC
    return d, g
