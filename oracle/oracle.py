"""ORACLE — test infrastructure only.

ctypes binding of oracle/_build/liboracle.so, the FP64 CPU restatement of the
AIFMM_LS 2D Lippmann-Schwinger path.  Imported only by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
checker, never as the measured or shipped path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

SCHED_EFFICIENT, SCHED_EAGER, SCHED_COLORED = 0, 1, 2


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.orc_create.restype = P
        L.orc_create.argtypes = [ip, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
        L.orc_destroy.argtypes = [P]
        L.orc_set_perturb.argtypes = [P, ctypes.c_double, ctypes.c_ulonglong]
        L.orc_set_find_phi.argtypes = [P, ctypes.c_int]
        L.orc_set_piv_eps.argtypes = [P, ctypes.c_double, ctypes.c_double]
        L.orc_zero_pivots.restype = ctypes.c_int
        L.orc_zero_pivots.argtypes = [P, ip, ctypes.c_int]
        L.orc_recompress_trace.restype = ctypes.c_int
        L.orc_recompress_trace.argtypes = [P, ip, ctypes.c_int]
        L.orc_n.restype = ctypes.c_longlong
        L.orc_n.argtypes = [P]
        L.orc_points.argtypes = [P, dp, dp]
        L.orc_tables.argtypes = [P, dp, dp, dp]
        L.orc_table.restype = ctypes.c_int
        L.orc_table.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.orc_entries.argtypes = [P, ctypes.c_longlong, lp, lp, dp]
        L.orc_bessel.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp]
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, dp, dp]
        L.orc_rhs.argtypes = [P, ctypes.c_double, dp]
        L.orc_solve.restype = ctypes.c_int
        L.orc_solve.argtypes = [P, dp, dp, ctypes.c_int, dp]
        L.orc_ranks.argtypes = [P, ip]
        L.orc_skeleton.restype = ctypes.c_int
        L.orc_skeleton.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
        L.orc_nca.argtypes = [P]
        L.orc_aca.restype = ctypes.c_int
        L.orc_aca.argtypes = [P, ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip, ip, ip, dp, dp]
        L.orc_dense.argtypes = [P, dp]
        L.orc_cpqr.restype = ctypes.c_int
        L.orc_cpqr.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_int, dp, dp, ip]
        L.orc_qr_solve.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp, dp]
        L.orc_lu_solve.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp, dp]
        L.orc_lists.argtypes = [P, ctypes.c_int, ctypes.c_int, ip]
        L.orc_adaptive_create.restype = P
        L.orc_adaptive_create.argtypes = [ip, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_int]
        L.orc_adaptive_destroy.argtypes = [P]
        L.orc_adaptive_n.restype = ctypes.c_longlong
        L.orc_adaptive_n.argtypes = [P]
        L.orc_adaptive_depth.restype = ctypes.c_int
        L.orc_adaptive_depth.argtypes = [P]
        L.orc_adaptive_splits.restype = ctypes.c_int
        L.orc_adaptive_splits.argtypes = [P]
        L.orc_adaptive_leaves.restype = ctypes.c_int
        L.orc_adaptive_leaves.argtypes = [P, ip, lp, dp, ctypes.c_int]
        L.orc_adaptive_points.argtypes = [P, dp, dp]
        L.orc_adaptive_neighbor.argtypes = [P, ctypes.c_int, ip, ip, ip, ip]
        L.orc_adaptive_set_find_phi.argtypes = [P, ctypes.c_int]
        L.orc_adaptive_entries.restype = ctypes.c_int
        L.orc_adaptive_entries.argtypes = [P, ctypes.c_longlong, lp, lp, dp]
        L.orc_adaptive_dense_matvec.restype = ctypes.c_int
        L.orc_adaptive_dense_matvec.argtypes = [P, dp, dp]
        _lib = L
    return _lib


def _c(z):
    return np.ascontiguousarray(np.asarray(z, dtype=np.complex128)).view(np.float64)


class Problem:
    """One discretised LS problem (tables, points, contrast) in the oracle."""

    def __init__(self, n_levels, kappa, tol_pow=8, qchoice=0, n_cheb=6, degree=6, L=1.0,
                 bumps=None, quantize=True, exact_bessel=False, perturb=0.0, perturb_seed=0, piv_eps=0.0,
                 base_piv_eps=0.0):
        ip = np.array([n_cheb, degree, n_levels, tol_pow, qchoice, int(quantize), int(exact_bessel)], np.int32)
        nb = 0
        self._bumps = None
        if bumps is not None:
            self._bumps = np.ascontiguousarray(np.asarray(bumps, np.float64).reshape(-1, 4))
            nb = self._bumps.shape[0]
        ptr = self._bumps.ctypes.data if self._bumps is not None else None
        self.h = lib().orc_create(ip, float(kappa), float(L), ptr, nb)
        self.n_levels, self.kappa, self.tol_pow, self.n_cheb, self.degree = n_levels, kappa, tol_pow, n_cheb, degree
        self.rank = n_cheb * n_cheb
        self.nbasis = degree * (degree + 1) // 2
        self.N = int(lib().orc_n(self.h))
        if perturb:
            # sensitivity probe: far-field entries scaled by 1 + perturb * u(i, j)
            lib().orc_set_perturb(self.h, float(perturb), int(perturb_seed))
        if piv_eps or base_piv_eps:
            # zero-pivot rule of the L2M_f factorisations and of the base system
            # (0: Eigen's, the reference's)
            lib().orc_set_piv_eps(self.h, float(piv_eps), float(base_piv_eps))

    def zero_pivots(self):
        """L2M_f factorisations of the last solve whose nonzero-pivot count fell
        below their size, one row each: level, box, nonzero pivots, size."""
        n = lib().orc_zero_pivots(self.h, np.zeros(1, np.int32), 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_zero_pivots(self.h, out, n)
        return out[:n].reshape(-1, 4).copy()

    def set_find_phi(self, on):
        """LS2DTree::findPhi for later entries() / dense(): True (default) the LS
        operator I + kappa^2 q G, False the volume potential G alone."""
        lib().orc_set_find_phi(self.h, int(bool(on)))

    def incident(self, theta=0.0):
        """The incident plane wave exp(i kappa d.x) at the points, d = (cos th,
        sin th) (IncidenceFunction, LS_functions.hpp:113-116, at th = 0)."""
        xy, _ = self.points()
        k = self.kappa
        if theta == 0.0:
            return np.exp(1j * k * xy[:, 0])
        return np.exp(1j * k * (np.cos(theta) * xy[:, 0] + np.sin(theta) * xy[:, 1]))

    def recompress_trace(self):
        """Recompressions of the last colour-schedule solve, one row each:
        level, colour step, box, row-stack rank, column-stack rank, new rank."""
        n = lib().orc_recompress_trace(self.h, np.zeros(1, np.int32), 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_recompress_trace(self.h, out, n)
        return out[:n].reshape(-1, 6).copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def points(self):
        xy = np.zeros(2 * self.N)
        q = np.zeros(self.N)
        lib().orc_points(self.h, xy, q)
        return xy.reshape(-1, 2), q

    def tables(self):
        M = np.zeros(121 * self.rank * 2)
        T = np.zeros(9 * self.rank * self.rank * 2)
        Qp = np.zeros(self.nbasis * self.rank * 2)
        lib().orc_tables(self.h, M, T, Qp)
        M = M.view(np.complex128).reshape(self.rank, 121).T
        T = T.view(np.complex128).reshape(9, self.rank, self.rank).transpose(0, 2, 1)
        Qp = Qp.view(np.complex128).reshape(self.rank, self.nbasis).T
        return M, T, Qp

    def table(self, family, level, n=0):
        """One precomputation table, unrounded: family -1 the far M at level
        (121 x rank); 0..3 (colleague, fine, separated fine, coarse) the
        neighbour operator n of a target box at level (rank x rank)."""
        rows = 121 if family < 0 else self.rank
        out = np.zeros(rows * self.rank * 2)
        if lib().orc_table(self.h, family, level, n, out) != rows:
            raise ValueError(f"no table family={family} level={level} n={n}")
        return out.view(np.complex128).reshape(self.rank, rows).T

    def entries(self, I, J):
        I = np.ascontiguousarray(I, np.int64)
        J = np.ascontiguousarray(J, np.int64)
        out = np.zeros(2 * I.size)
        lib().orc_entries(self.h, I.size, I, J, out)
        return out.view(np.complex128)

    def rhs(self, theta=0.0):
        b = np.zeros(2 * self.N)
        lib().orc_rhs(self.h, float(theta), b)
        return b.view(np.complex128)

    def solve(self, b, schedule=SCHED_EFFICIENT):
        x = np.zeros(2 * self.N)
        st = np.zeros(8)
        lib().orc_solve(self.h, _c(b), x, int(schedule), st)
        return x.view(np.complex128), st

    def dense(self):
        A = np.zeros(2 * self.N * self.N)
        lib().orc_dense(self.h, A)
        return A.view(np.complex128).reshape(self.N, self.N).T

    def nca(self):
        lib().orc_nca(self.h)

    def ranks(self):
        nl = 4 ** self.n_levels
        out = np.zeros((self.n_levels + 1) * nl, np.int32)
        lib().orc_ranks(self.h, out)
        return {j: out[j * nl: j * nl + 4 ** j].copy() for j in range(2, self.n_levels + 1)}

    def skeleton(self, j, k, which):
        out = np.zeros(4 * self.N, np.int32)
        n = lib().orc_skeleton(self.h, j, k, which, out)
        return out[:n].copy()

    def lists(self, level, box):
        out = np.zeros(48, np.int32)
        lib().orc_lists(self.h, level, box, out)
        return out[:8].copy(), out[8:24].copy(), out[24:].copy()

    def aca(self, rows, cols, tol_pow):
        rows = np.ascontiguousarray(rows, np.int32)
        cols = np.ascontiguousarray(cols, np.int32)
        mn = min(rows.size, cols.size)
        rb = np.zeros(mn, np.int32)
        cb = np.zeros(mn, np.int32)
        Ac = np.zeros(2 * rows.size * mn)
        Ar = np.zeros(2 * cols.size * mn)
        r = lib().orc_aca(self.h, tol_pow, rows.size, rows, cols.size, cols, rb, cb, Ac, Ar)
        Ac = Ac[: 2 * rows.size * r].view(np.complex128).reshape(r, rows.size).T
        Ar = Ar[: 2 * cols.size * r].view(np.complex128).reshape(cols.size, r).T
        return rb[:r].copy(), cb[:r].copy(), Ac, Ar


def bessel(x, nmax=60, exact=False):
    J = np.zeros(nmax + 1)
    Y = np.zeros(nmax + 1)
    lib().orc_bessel(float(x), nmax, int(exact), J, Y)
    return J, Y


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, x, w)
    return x, w


def cpqr(A, thr, k):
    A = np.asarray(A, np.complex128)
    m, n = A.shape
    Q = np.zeros(2 * m * k)
    rd = np.zeros(min(m, n))
    perm = np.zeros(n, np.int32)
    r = lib().orc_cpqr(m, n, _c(A.T.copy()), float(thr), k, Q, rd, perm)
    return r, Q.view(np.complex128).reshape(k, m).T, rd, perm


def qr_solve(A, B):
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128).reshape(A.shape[0], -1)
    n, nr = A.shape[0], B.shape[1]
    X = np.zeros(2 * n * nr)
    lib().orc_qr_solve(n, _c(A.T.copy()), nr, _c(B.T.copy()), X)
    return X.view(np.complex128).reshape(nr, n).T


def lu_solve(A, B):
    A = np.asarray(A, np.complex128)
    B = np.asarray(B, np.complex128).reshape(A.shape[0], -1)
    n, nr = A.shape[0], B.shape[1]
    X = np.zeros(2 * n * nr)
    lib().orc_lu_solve(n, _c(A.T.copy()), nr, _c(B.T.copy()), X)
    return X.view(np.complex128).reshape(nr, n).T


class AdaptiveProblem:
    """The adaptive quadtree with 2:1 level restriction and the entries
    between its leaves (LS2DTree::createAdaptiveTree, makeLevelRestriction_Tree,
    getIfNeighbor, getMatrixEntry) in the oracle."""

    def __init__(self, kappa, tree_adaptivity=1, qchoice=0, n_cheb=6, degree=6, L=1.0, bumps=None, quantize=True,
                 exact_bessel=False):
        ip = np.array([n_cheb, degree, 0, 8, qchoice, int(quantize), int(exact_bessel)], np.int32)
        nb = 0
        self._bumps = None
        if bumps is not None:
            self._bumps = np.ascontiguousarray(np.asarray(bumps, np.float64).reshape(-1, 4))
            nb = self._bumps.shape[0]
        ptr = self._bumps.ctypes.data if self._bumps is not None else None
        self.h = lib().orc_adaptive_create(ip, float(kappa), float(L), ptr, nb, int(tree_adaptivity))
        self.kappa, self.L, self.rank = kappa, L, n_cheb * n_cheb
        self.N = int(lib().orc_adaptive_n(self.h))
        self.depth = int(lib().orc_adaptive_depth(self.h))
        self.restriction_splits = int(lib().orc_adaptive_splits(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_adaptive_destroy(self.h)
            self.h = None

    def leaves(self):
        """(levels, box numbers, centres) in leafNodes order."""
        n = self.N // self.rank
        lv, bx, c = np.zeros(n, np.int32), np.zeros(n, np.int64), np.zeros(2 * n)
        lib().orc_adaptive_leaves(self.h, lv, bx, c, n)
        return lv, bx, c.reshape(-1, 2)

    def points(self):
        xy = np.zeros(2 * self.N)
        q = np.zeros(self.N)
        lib().orc_adaptive_points(self.h, xy, q)
        return xy.reshape(-1, 2), q

    def neighbor(self, li, lj):
        """getIfNeighbor of leaf pairs: (type, slot); type -1 far, 0 colleague,
        1 coarse, 2 fine, 3 separated fine."""
        li = np.ascontiguousarray(li, np.int32)
        lj = np.ascontiguousarray(lj, np.int32)
        t, s = np.zeros(li.size, np.int32), np.zeros(li.size, np.int32)
        lib().orc_adaptive_neighbor(self.h, li.size, li, lj, t, s)
        return t, s

    def set_find_phi(self, on):
        lib().orc_adaptive_set_find_phi(self.h, int(bool(on)))

    def entries(self, I, J):
        I = np.ascontiguousarray(I, np.int64)
        J = np.ascontiguousarray(J, np.int64)
        out = np.zeros(2 * I.size)
        if lib().orc_adaptive_entries(self.h, I.size, I, J, out):
            raise RuntimeError("adaptive oracle: a near pair matched no neighbour slot")
        return out.view(np.complex128)

    def dense_matvec(self, x):
        y = np.zeros(2 * self.N)
        if lib().orc_adaptive_dense_matvec(self.h, _c(x), y):
            raise RuntimeError("adaptive oracle: a near pair matched no neighbour slot")
        return y.view(np.complex128)
