"""MI355X-native 2D Lippmann-Schwinger direct solver (AIFMM_LS solve path).

Thin ctypes binding of the in-tree C ABI library ``_lib/libaifmm_ls.so``
(declared in ``include/aifmm_ls.h``).  All arithmetic runs in hand-written
gfx950 HIP kernels; there is no CPU fallback: if the library or a GPU is
missing, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libaifmm_ls.so")
_lib = None

PREC_C128, PREC_C64 = 0, 1


class AifmmError(RuntimeError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [("n_cheb", ctypes.c_int), ("degree", ctypes.c_int), ("n_levels", ctypes.c_int),
                ("tol_pow", ctypes.c_int), ("qchoice", ctypes.c_int), ("quantize_tables", ctypes.c_int),
                ("kappa", ctypes.c_double), ("L", ctypes.c_double),
                ("bumps", ctypes.POINTER(ctypes.c_double)), ("n_bumps", ctypes.c_int)]


def lib():
    """Load the C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AifmmError(f"{LIB_PATH} missing: run `python -m aifmm_ls_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        L.aifmm_ls_create.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_void_p)]
        L.aifmm_ls_create_cached.argtypes = [ctypes.POINTER(_Params), ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
        L.aifmm_ls_write_precomputations.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
        L.aifmm_ls_create_adaptive.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.aifmm_ls_leaves.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64), dp,
                                      ctypes.POINTER(ctypes.c_int)]
        L.aifmm_ls_table.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.aifmm_ls_destroy.argtypes = [P]
        L.aifmm_ls_last_error.restype = ctypes.c_char_p
        L.aifmm_ls_size.restype = ctypes.c_int64
        L.aifmm_ls_size.argtypes = [P]
        L.aifmm_ls_points.argtypes = [P, dp, dp]
        L.aifmm_ls_tables.argtypes = [P, dp, dp, ctypes.POINTER(ctypes.c_int)]
        L.aifmm_ls_entries.argtypes = [P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_int64), dp]
        L.aifmm_ls_rhs_planewave.argtypes = [P, ctypes.c_double, dp]
        L.aifmm_ls_assemble.argtypes = [P]
        L.aifmm_ls_ranks.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.aifmm_ls_skeleton.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]
        L.aifmm_ls_factor.argtypes = [P, ctypes.c_int]
        L.aifmm_ls_factor_batch.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, dp]
        L.aifmm_ls_set_pivot_eps.argtypes = [P, ctypes.c_double, ctypes.c_double]
        L.aifmm_ls_set_growth_gate.argtypes = [P, ctypes.c_double]
        L.aifmm_ls_factor_trace.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int)]
        L.aifmm_ls_solve.argtypes = [P, ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_int, dp]
        L.aifmm_ls_set_rhs.argtypes = [P, ctypes.c_int, dp]
        L.aifmm_ls_solve_resident.argtypes = [P, ctypes.c_double, ctypes.c_int, dp]
        L.aifmm_ls_get_solution.argtypes = [P, dp]
        L.aifmm_ls_stats.argtypes = [P, dp, ctypes.c_int]
        L.aifmm_ls_matvec.argtypes = [P, ctypes.c_int, dp, dp]
        L.aifmm_ls_dense_matvec.argtypes = [P, ctypes.c_int, dp, dp]
        L.aifmm_ls_set_find_phi.argtypes = [P, ctypes.c_int]
        L.aifmm_ls_incident_field.argtypes = [P, ctypes.c_double, dp]
        L.aifmm_ls_total_field.argtypes = [P, ctypes.c_double, dp, dp]
        L.aifmm_ls_dense_matvec_noalloc.argtypes = [P, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p]
        ip64 = ctypes.POINTER(ctypes.c_int64)
        ip = ctypes.POINTER(ctypes.c_int)
        L.aifmm_ls_aca.argtypes = [P, ctypes.c_int, ctypes.c_int, ip64, ctypes.c_int, ip64, ip, ip, dp, dp, ip]
        L.aifmm_ls_selftest.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp,
                                        ctypes.POINTER(ctypes.c_int), ctypes.c_double]
        L.aifmm_ls_profile_enable.argtypes = [ctypes.c_int]
        L.aifmm_ls_selftest_stream.argtypes = [ctypes.c_int64, ctypes.c_int, dp]
        L.aifmm_ls_set_refinement.argtypes = [P, ctypes.c_int]
        L.aifmm_ls_profile_read.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_char_p, ctypes.c_int]
        L.aifmm_ls_profile_launches.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_char_p, ctypes.c_int]
        L.aifmm_ls_set_tuning.argtypes = [ctypes.c_int]
        L.aifmm_ls_selftest_prog.argtypes = [ctypes.c_int] * 7 + [dp] * 6
        L.aifmm_ls_device_bytes.argtypes = [P, ip64]
        L.aifmm_ls_solve_batch.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                           ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.c_double, ctypes.c_int,
                                           dp, dp]
        L.aifmm_ls_server_create.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.aifmm_ls_server_factor.argtypes = [P, ctypes.POINTER(_Params), ctypes.c_int, ip64, ip64]
        L.aifmm_ls_server_solve.argtypes = [P, ctypes.c_int64, ctypes.c_int, dp, dp, ctypes.c_double, ctypes.c_int,
                                            ip64]
        L.aifmm_ls_server_wait.argtypes = [P, ctypes.c_int64, dp]
        L.aifmm_ls_server_release.argtypes = [P, ctypes.c_int64]
        L.aifmm_ls_server_stats.argtypes = [P, dp, ctypes.c_int]
        L.aifmm_ls_server_destroy.argtypes = [P]
        _lib = L
    return _lib


def _chk(rc):
    if rc != 0:
        raise AifmmError(lib().aifmm_ls_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def profile_enable(on=True, launches=False):
    """Record HIP event pairs around every batched launch (see profile_read);
    launches=True also keeps one record per launch (see profile_launches)."""
    _chk(lib().aifmm_ls_profile_enable((2 if launches else 1) if on else 0))


def set_tuning(mode=1):
    """Launch tuning for programs not yet tuned: 0/False off, 1/True timed,
    or 256 / 512 / 1024 threads per workgroup wherever they fit (see
    aifmm_ls_set_tuning)."""
    _chk(lib().aifmm_ls_set_tuning(int(mode)))


def profile_launches(reset=True, max_launches=100000):
    """Per-launch records since the last reset, in launch order: list of
    (kernel name, algorithmic bytes, milliseconds)."""
    out = np.zeros(2 * max_launches)
    names = ctypes.create_string_buffer(48 * max_launches)
    n = lib().aifmm_ls_profile_launches(1 if reset else 0, max_launches, _dp(out), names, len(names))
    if n < 0:
        raise AifmmError(lib().aifmm_ls_last_error().decode())
    nm = names.value.decode().split("\n")
    return [(nm[i], out[2 * i], out[2 * i + 1]) for i in range(n)]


def profile_read(reset=True, max_kernels=64):
    """Per-kernel totals since the last reset: list of dicts with name,
    launches, ms (summed event durations), bytes and flops (algorithmic)."""
    out = np.zeros(5 * max_kernels)
    names = ctypes.create_string_buffer(64 * max_kernels)
    n = lib().aifmm_ls_profile_read(1 if reset else 0, max_kernels, _dp(out), names, len(names))
    if n < 0:
        raise AifmmError(lib().aifmm_ls_last_error().decode())
    nm = names.value.decode().split("\n")
    return [dict(name=nm[i], launches=int(out[5 * i]), ms=out[5 * i + 1], bytes=out[5 * i + 2],
                 flops=out[5 * i + 3]) for i in range(n)]


STAT_NAMES = ["t_tables", "t_assemble", "t_factor", "factor_flops", "factor_bytes", "base_n", "max_rank",
              "solve_bytes_1rhs", "recompressed_boxes", "operator_bytes", "t_solve", "t_nca", "ill_boxes",
              "sweep_tuned_time_ratio", "sweep_tuned_stages", "dense_top_level", "dense_top_n",
              "base_rdiag_ratio", "zeroed_pivot_blocks", "pivot_blocks_c64", "pivot_blocks_fp64"]


class LSProblem:
    """The reference's userkernel + FMM2DTree pipeline on one GPU.

    Arguments follow the reference driver's argv (testFMM2D_find_error.cpp:44-52):
    n_cheb, kappa, degree, L, qchoice, tol_pow, n_levels; qchoice 6 takes
    ``bumps`` = rows of (a, w, cx, cy).  ``precomputations`` = a directory
    holding the reference's precomputations/ text cache: the tables are read
    from it instead of computed (LS_Assembly.hpp:1217-1228, 1422-1459; the
    reference with its three precomputation lines commented out, LS/Readme.md).
    ``tree_adaptivity`` = t: the adaptive quadtree with 2:1 level restriction
    instead of the uniform one, refined until the contrast interpolates to
    10^-t (aifmm_ls_create_adaptive; n_levels unused; entries, the dense
    product and the fields only).
    """

    def __init__(self, n_levels, kappa, tol_pow=8, qchoice=0, n_cheb=6, degree=6, L=1.0, bumps=None,
                 quantize=True, precomputations=None, tree_adaptivity=None):
        p = _Params()
        p.n_cheb, p.degree, p.n_levels, p.tol_pow, p.qchoice = n_cheb, degree, n_levels, tol_pow, qchoice
        p.quantize_tables = int(quantize)
        p.kappa, p.L = float(kappa), float(L)
        self._bumps = None
        if bumps is not None:
            self._bumps = np.ascontiguousarray(np.asarray(bumps, np.float64).reshape(-1, 4))
            p.bumps = _dp(self._bumps)
            p.n_bumps = self._bumps.shape[0]
        h = ctypes.c_void_p()
        if tree_adaptivity is not None:
            # the adaptive quadtree (n_levels unused): aifmm_ls_create_adaptive
            _chk(lib().aifmm_ls_create_adaptive(ctypes.byref(p), int(tree_adaptivity), ctypes.byref(h)))
        elif precomputations is None:
            _chk(lib().aifmm_ls_create(ctypes.byref(p), ctypes.byref(h)))
        else:
            _chk(lib().aifmm_ls_create_cached(ctypes.byref(p), os.fsencode(precomputations), ctypes.byref(h)))
        self.h = h
        self.N = int(lib().aifmm_ls_size(h))
        self.n_levels, self.kappa, self.tol_pow = n_levels, kappa, tol_pow
        self.R = n_cheb * n_cheb

    def close(self):
        if getattr(self, "h", None):
            lib().aifmm_ls_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def leaves(self):
        """The adaptive tree's leaves in order: (levels, box numbers, centres)."""
        n = ctypes.c_int(0)
        _chk(lib().aifmm_ls_leaves(self.h, None, None, None, ctypes.byref(n)))
        lv = np.zeros(n.value, np.int32)
        bx = np.zeros(n.value, np.int64)
        c = np.zeros(2 * n.value)
        _chk(lib().aifmm_ls_leaves(self.h, lv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   bx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(c), ctypes.byref(n)))
        return lv, bx, c.reshape(-1, 2)

    def points(self):
        xy = np.zeros(2 * self.N)
        q = np.zeros(self.N)
        _chk(lib().aifmm_ls_points(self.h, _dp(xy), _dp(q)))
        return xy.reshape(-1, 2), q

    def write_precomputations(self, directory):
        """Every table of the reference's precomputation (far M at levels
        0..10; colleague, fine, separated-fine and coarse neighbour operators
        at levels 2..n_levels) computed on the device and written as the
        reference's text cache under directory/precomputations
        (LS_Assembly.hpp:1166-1184, 1328-1420).  Returns the device seconds."""
        t = ctypes.c_double()
        _chk(lib().aifmm_ls_write_precomputations(self.h, os.fsencode(directory), ctypes.byref(t)))
        return t.value

    def table(self, family, level, n=0):
        """One precomputation table, unrounded: family -1 the far M at level
        (121 x R); 0..3 (colleague, fine, separated fine, coarse) the neighbour
        operator n of a target box at level (R x R)."""
        rows = 121 if family < 0 else self.R
        out = np.zeros(2 * rows * self.R)
        _chk(lib().aifmm_ls_table(self.h, family, level, n, _dp(out)))
        return out.view(np.complex128).reshape(self.R, rows).T

    def tables(self):
        M = np.zeros(2 * 121 * self.R)
        T = np.zeros(2 * 9 * self.R * self.R)
        c = ctypes.c_int()
        _chk(lib().aifmm_ls_tables(self.h, _dp(M), _dp(T), ctypes.byref(c)))
        M = M.view(np.complex128).reshape(self.R, 121).T
        T = T.view(np.complex128).reshape(9, self.R, self.R).transpose(0, 2, 1)
        return M, T, c.value

    def entries(self, I, J):
        I = np.ascontiguousarray(I, np.int64)
        J = np.ascontiguousarray(J, np.int64)
        out = np.zeros(2 * I.size)
        _chk(lib().aifmm_ls_entries(self.h, I.size, I.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                    J.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(out)))
        return out.view(np.complex128)

    def rhs(self, theta=0.0):
        b = np.zeros(2 * self.N)
        _chk(lib().aifmm_ls_rhs_planewave(self.h, float(theta), _dp(b)))
        return b.view(np.complex128)

    def assemble(self):
        _chk(lib().aifmm_ls_assemble(self.h))

    def ranks(self, level):
        out = np.zeros(4 ** level, np.int32)
        _chk(lib().aifmm_ls_ranks(self.h, level, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def skeleton(self, level, box, which):
        out = np.zeros(4 * 36 * 64, np.int64)
        n = ctypes.c_int(out.size)
        _chk(lib().aifmm_ls_skeleton(self.h, level, box, which, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     ctypes.byref(n)))
        return out[:n.value].copy()

    def factor(self, precision=PREC_C128):
        _chk(lib().aifmm_ls_factor(self.h, precision))

    def set_pivot_eps(self, l2mf_eps=1e-12, base_eps=0.0):
        """Zero-pivot epsilons of the next complex128 factorisation: the L2M_f
        pivot blocks (default 1e-12) and the level-2 base system (default 0 =
        Eigen's rule, the reference's)."""
        _chk(lib().aifmm_ls_set_pivot_eps(self.h, float(l2mf_eps), float(base_eps)))

    def set_growth_gate(self, thr=1e4):
        """Pivot blocks of the next complex64 factorisation: thr < 0 factors
        all of them in FP64; thr >= 0 (default 1e4) keeps a block's complex64
        LU when its growth factor times its condition estimate is at most thr."""
        _chk(lib().aifmm_ls_set_growth_gate(self.h, float(thr)))

    def factor_trace(self, which=0):
        """which 0: recompressions of the factorisation (rows: level, colour
        step, box, row-stack rank, column-stack rank, new rank); which 1:
        L2M_f factorisations with zeroed pivots (rows: level, box, nonzero
        pivots, size; complex128 only)."""
        n = ctypes.c_int()
        ip = ctypes.POINTER(ctypes.c_int)
        _chk(lib().aifmm_ls_factor_trace(self.h, which, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), np.int32)
        _chk(lib().aifmm_ls_factor_trace(self.h, which, out.ctypes.data_as(ip), n.value, ctypes.byref(n)))
        return out[:n.value].reshape(-1, 6 if which == 0 else 4).copy()

    def solve(self, b, tol=1e-6, max_refine=0):
        b = np.asarray(b, np.complex128)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        bb = np.ascontiguousarray(b.reshape(self.N, nrhs).T).view(np.float64)  # column-major N x nrhs
        x = np.zeros(2 * self.N * nrhs)
        st = np.zeros(8)
        _chk(lib().aifmm_ls_solve(self.h, nrhs, _dp(bb), _dp(x), tol, max_refine, _dp(st)))
        x = x.view(np.complex128).reshape(nrhs, self.N).T
        return (x[:, 0] if b.ndim == 1 else x), st

    def set_rhs(self, b):
        b = np.asarray(b, np.complex128)
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        bb = np.ascontiguousarray(b.reshape(self.N, nrhs).T).view(np.float64)
        _chk(lib().aifmm_ls_set_rhs(self.h, nrhs, _dp(bb)))
        self._nrhs = nrhs

    def solve_resident(self, tol=1e-6, max_refine=0):
        st = np.zeros(8)
        _chk(lib().aifmm_ls_solve_resident(self.h, tol, max_refine, _dp(st)))
        return st

    def set_refinement(self, method):
        """0: flexible GMRES with the factorisation as preconditioner (default);
        1: plain iterative refinement."""
        _chk(lib().aifmm_ls_set_refinement(self.h, int(method)))

    def get_solution(self):
        x = np.zeros(2 * self.N * self._nrhs)
        _chk(lib().aifmm_ls_get_solution(self.h, _dp(x)))
        return x.view(np.complex128).reshape(self._nrhs, self.N).T

    def matvec(self, x):
        x = np.asarray(x, np.complex128)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        xx = np.ascontiguousarray(x.reshape(self.N, nrhs).T).view(np.float64)
        y = np.zeros(2 * self.N * nrhs)
        _chk(lib().aifmm_ls_matvec(self.h, nrhs, _dp(xx), _dp(y)))
        y = y.view(np.complex128).reshape(nrhs, self.N).T
        return y[:, 0] if x.ndim == 1 else y

    def dense_matvec(self, x):
        """y = A x with the discretised operator itself (every entry as
        getMatrixEntry), as the generator's dense rhs = A x."""
        x = np.asarray(x, np.complex128)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        xx = np.ascontiguousarray(x.reshape(self.N, nrhs).T).view(np.float64)
        y = np.zeros(2 * self.N * nrhs)
        _chk(lib().aifmm_ls_dense_matvec(self.h, nrhs, _dp(xx), _dp(y)))
        y = y.view(np.complex128).reshape(nrhs, self.N).T
        return y[:, 0] if x.ndim == 1 else y

    def set_find_phi(self, on):
        """LS2DTree::findPhi for entries() and dense_matvec(): True (default)
        the LS operator I + kappa^2 q G, False the volume potential G."""
        _chk(lib().aifmm_ls_set_find_phi(self.h, int(bool(on))))

    def incident_field(self, theta=0.0):
        """exp(i kappa d.x) at the points, d = (cos theta, sin theta)."""
        u = np.zeros(2 * self.N)
        _chk(lib().aifmm_ls_incident_field(self.h, float(theta), _dp(u)))
        return u.view(np.complex128)

    def scattered_field(self, phi):
        """G phi with the dense volume-potential product (findPhi false)."""
        self.set_find_phi(False)
        try:
            return self.dense_matvec(phi)
        finally:
            self.set_find_phi(True)

    def total_field(self, phi, theta=0.0):
        """u_inc + G phi for the plane wave at theta (hybrid.cpp:264-281)."""
        phi = np.ascontiguousarray(np.asarray(phi, np.complex128))
        u = np.zeros(2 * self.N)
        _chk(lib().aifmm_ls_total_field(self.h, float(theta), _dp(phi.view(np.float64)), _dp(u)))
        return u.view(np.complex128)

    def aca(self, rows, cols, tol_pow):
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        mn = min(rows.size, cols.size)
        rb = np.zeros(max(mn, 1), np.int32)
        cb = np.zeros(max(mn, 1), np.int32)
        Ac = np.zeros(2 * rows.size * max(mn, 1))
        Ar = np.zeros(2 * cols.size * max(mn, 1))
        r = ctypes.c_int()
        ip = ctypes.POINTER(ctypes.c_int)
        _chk(lib().aifmm_ls_aca(self.h, tol_pow, rows.size, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                cols.size, cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                rb.ctypes.data_as(ip), cb.ctypes.data_as(ip), _dp(Ac), _dp(Ar), ctypes.byref(r)))
        r = r.value
        Ac = Ac[: 2 * rows.size * r].view(np.complex128).reshape(r, rows.size).T
        Ar = Ar[: 2 * cols.size * r].view(np.complex128).reshape(cols.size, r).T
        return rb[:r].copy(), cb[:r].copy(), Ac, Ar

    def device_bytes(self):
        """Device memory the problem holds (operator, factors, vectors)."""
        b = ctypes.c_int64()
        _chk(lib().aifmm_ls_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def stats(self):
        out = np.zeros(len(STAT_NAMES))
        _chk(lib().aifmm_ls_stats(self.h, _dp(out), len(STAT_NAMES)))
        return dict(zip(STAT_NAMES, out.tolist()))


def _cm(a):
    return np.ascontiguousarray(np.asarray(a, np.complex128).T).view(np.float64)


def selftest_stream(nbytes, reps=3):
    """Copy nbytes between two fresh device buffers reps times with the plain
    copy kernel (k_stream_copy); returns the mean milliseconds per copy."""
    ms = ctypes.c_double()
    _chk(lib().aifmm_ls_selftest_stream(int(nbytes), int(reps), ctypes.byref(ms)))
    return ms.value


def selftest_gemm(A, B, c64=False):
    m, k = A.shape
    n = B.shape[1]
    C = np.zeros(2 * m * n)
    _chk(lib().aifmm_ls_selftest(10 * c64 + 0, m, n, k, _dp(_cm(A)), _dp(_cm(B)), _dp(C), None, 0.0))
    return C.view(np.complex128).reshape(n, m).T


def selftest_lu_solve(A, B, c64=False):
    m, n = B.shape
    C = np.zeros(2 * m * n)
    _chk(lib().aifmm_ls_selftest(10 * c64 + 1, m, n, 0, _dp(_cm(A)), _dp(_cm(B)), _dp(C), None, 0.0))
    return C.view(np.complex128).reshape(n, m).T


def selftest_gated_lu_solve(A, B, thr):
    """A^{-1} B through the complex64 pivot-block growth gate at threshold thr;
    returns (X, estimate, went_to_fp64)."""
    m, n = B.shape
    C = np.zeros(2 * m * n + 2)
    fl = (ctypes.c_int * 1)()
    _chk(lib().aifmm_ls_selftest(14, m, n, 0, _dp(_cm(A)), _dp(_cm(B)), _dp(C), fl, float(thr)))
    return C[:2 * m * n].view(np.complex128).reshape(n, m).T, float(C[2 * m * n]), bool(fl[0])


def selftest_qr_solve(A, B, c64=False):
    """A^{-1} B by the batched column-pivoted QR kernels; returns (X, nonzero pivots)."""
    m, n = B.shape
    C = np.zeros(2 * m * n)
    nz = (ctypes.c_int * 1)()
    _chk(lib().aifmm_ls_selftest(10 * c64 + 3, m, n, 0, _dp(_cm(A)), _dp(_cm(B)), _dp(C), nz, 0.0))
    return C.view(np.complex128).reshape(n, m).T, nz[0]


def selftest_cpqr(A, thr, k, c64=False):
    m, n = A.shape
    C = np.zeros(2 * m * k)
    r = ctypes.c_int()
    dummy = np.zeros(2)
    _chk(lib().aifmm_ls_selftest(10 * c64 + 2, m, n, k, _dp(_cm(A)), _dp(dummy), _dp(C), ctypes.byref(r), thr))
    return r.value, C.view(np.complex128).reshape(k, m).T


def selftest_cpqr_raw(A, thr, reg, early=False, c64=False):
    """Column-pivoted QR of A by one kernel (reg: the register kernel, else
    the global-memory one): (R and reflectors as one m x n array, tau,
    permutation, rank, nonzero pivots)."""
    m, n = A.shape
    C = np.zeros(2 * (m * n + min(m, n)))
    io = np.zeros(2 + n, dtype=np.int32)
    _chk(lib().aifmm_ls_selftest(10 * c64 + 5, m, n, int(bool(reg)) + 2 * int(bool(early)), _dp(_cm(A)),
                                 _dp(np.zeros(2)), _dp(C), io.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), thr))
    z = C.view(np.complex128)
    return z[:m * n].reshape(n, m).T, z[m * n:], io[2:].copy(), int(io[0]), int(io[1])


def _blocks(a):
    """Stacked blocks (..., r, c) to consecutive column-major interleaved complex."""
    return np.ascontiguousarray(np.swapaxes(np.asarray(a, np.complex128), -1, -2)).view(np.float64)


def selftest_program(A, X, S, Bm, shape=0, c64=False):
    """Box programs on the GPU (see aifmm_ls_selftest_prog): A (nprog, nterms,
    m, cols), X (nprog, nterms, cols, nrhs), S (nprog, m, nrhs), Bm (nprog, m,
    m); returns Y = S + sum_t alpha_t A X and Out = Bm Y, each (nprog, m, nrhs)."""
    nprog, nterms, m, cols = A.shape
    nrhs = X.shape[-1]
    if X.shape != (nprog, nterms, cols, nrhs) or S.shape != (nprog, m, nrhs) or Bm.shape != (nprog, m, m):
        raise AifmmError("self-test block shapes do not match")
    Y = np.zeros(2 * nprog * m * nrhs)
    O = np.zeros(2 * nprog * m * nrhs)
    _chk(lib().aifmm_ls_selftest_prog(1 if c64 else 0, int(shape), nprog, m, nterms, cols, nrhs, _dp(_blocks(A)),
                                      _dp(_blocks(X)), _dp(_blocks(S)), _dp(_blocks(Bm)), _dp(Y), _dp(O)))
    back = lambda v: np.swapaxes(v.view(np.complex128).reshape(nprog, nrhs, m), -1, -2)
    return back(Y), back(O)


def _params(n_levels, kappa, tol_pow, qchoice, n_cheb, degree, L, bumps, quantize):
    p = _Params()
    p.n_cheb, p.degree, p.n_levels, p.tol_pow, p.qchoice = n_cheb, degree, n_levels, tol_pow, qchoice
    p.quantize_tables = int(quantize)
    p.kappa, p.L = float(kappa), float(L)
    keep = None
    if bumps is not None:
        keep = np.ascontiguousarray(np.asarray(bumps, np.float64).reshape(-1, 4))
        p.bumps = _dp(keep)
        p.n_bumps = keep.shape[0]
    return p, keep


def factor_batch(problems, precision=PREC_C64):
    """Factorise a batch of LSProblems together (aifmm_ls_factor_batch): the
    eliminations of all of them as one merged level-synchronous launch
    sequence.  Returns {"wall_s", "factor_s", "rounds", "launches",
    "launch_s", "gather_s"}: wall seconds (assembly included), seconds of
    the factorisations, launch rounds, kernel launches, seconds spent
    launching, and seconds between rounds (the members' host work)."""
    n = len(problems)
    hs = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    info = np.zeros(6)
    _chk(lib().aifmm_ls_factor_batch(hs, n, int(precision), _dp(info)))
    return {"wall_s": float(info[0]), "factor_s": float(info[1]), "rounds": int(info[2]),
            "launches": int(info[3]), "launch_s": float(info[4]), "gather_s": float(info[5])}


def solve_batch(problems, rhs, tol=1e-9, max_refine=30):
    """Solve one right-hand side per problem for a batch of factorised
    LSProblems with merged level-synchronous launches (aifmm_ls_solve_batch).
    Returns (list of x, per-problem relative residuals, stats = [time, steps,
    max relres])."""
    n = len(problems)
    hs = (ctypes.c_void_p * n)(*[p.h.value for p in problems])
    bs = [np.ascontiguousarray(np.asarray(b, np.complex128)).view(np.float64) for b in rhs]
    xs = [np.zeros(2 * p.N) for p in problems]
    for p, b in zip(problems, bs):
        if b.size != 2 * p.N:
            raise AifmmError("right-hand side size does not match its problem")
    dp = ctypes.POINTER(ctypes.c_double)
    bp = (dp * n)(*[_dp(b) for b in bs])
    xp = (dp * n)(*[_dp(x) for x in xs])
    rel = np.zeros(n)
    st = np.zeros(3)
    _chk(lib().aifmm_ls_solve_batch(hs, n, bp, xp, float(tol), int(max_refine), _dp(rel), _dp(st)))
    return [x.view(np.complex128) for x in xs], rel, st


SERVER_STAT_NAMES = ["resident_problems", "resident_bytes", "budget", "evictions", "refactorisations", "completed",
                     "batches", "merged_requests", "peak_running", "queued", "cross_problem_batches"]


class Server:
    """The serving loop of the C ABI (aifmm_ls_server_*): factorised problems
    stay resident on the GPU, requests run on `slots` worker threads (each on
    its problem's own HIP stream, so different problems overlap on the
    device), queued solves of one problem are merged into one multi-RHS solve,
    and new factorisations are admitted by device memory against
    `budget_bytes` with least-recently-used eviction.

    ``factor(...)`` and ``solve(...)`` queue a request and return at once;
    ``wait(ticket)`` returns (x or None, info dict)."""

    def __init__(self, budget_bytes, slots=4):
        s = ctypes.c_void_p()
        _chk(lib().aifmm_ls_server_create(int(budget_bytes), int(slots), ctypes.byref(s)))
        self.s = s
        self._out = {}
        self._N = {}

    def factor(self, n_levels, kappa, tol_pow=8, qchoice=0, bumps=None, precision=PREC_C64, n_cheb=6, degree=6,
               L=1.0, quantize=True):
        """Queue a factorisation; returns (problem id, ticket)."""
        p, _keep = _params(n_levels, kappa, tol_pow, qchoice, n_cheb, degree, L, bumps, quantize)
        pid, t = ctypes.c_int64(), ctypes.c_int64()
        _chk(lib().aifmm_ls_server_factor(self.s, ctypes.byref(p), int(precision), ctypes.byref(pid),
                                          ctypes.byref(t)))
        self._N[pid.value] = (1 << (2 * n_levels)) * n_cheb * n_cheb
        self._out[t.value] = None
        return pid.value, t.value

    def solve(self, problem, b, tol=1e-9, max_refine=30):
        """Queue a solve of b (N or N x nrhs complex); returns the ticket."""
        b = np.asarray(b, np.complex128)
        vec = b.ndim == 1
        B = b.reshape(b.shape[0], -1)
        N = self._N[problem]
        if B.shape[0] != N:
            raise AifmmError(f"right-hand side has {B.shape[0]} rows, problem has {N}")
        nrhs = B.shape[1]
        bb = _cm(B)
        x = np.zeros(2 * N * nrhs)
        t = ctypes.c_int64()
        _chk(lib().aifmm_ls_server_solve(self.s, int(problem), nrhs, _dp(bb), _dp(x), float(tol), int(max_refine),
                                         ctypes.byref(t)))
        self._out[t.value] = (x, N, nrhs, vec)
        return t.value

    def wait(self, ticket):
        info = np.zeros(9)
        rc = lib().aifmm_ls_server_wait(self.s, int(ticket), _dp(info))
        out = self._out.pop(ticket, None)
        if rc != 0:
            raise AifmmError(lib().aifmm_ls_last_error().decode())
        d = dict(latency=info[1], queue=info[2], run=info[3], steps=int(info[4]), relres=info[5],
                 batch_rhs=int(info[6]), refactored=bool(info[7]), batch_problems=int(info[8]))
        if out is None:
            return None, d
        x, N, nrhs, vec = out
        X = x.view(np.complex128).reshape(nrhs, N).T
        return (X[:, 0].copy() if vec else X.copy()), d

    def release(self, problem):
        _chk(lib().aifmm_ls_server_release(self.s, int(problem)))

    def stats(self):
        out = np.zeros(len(SERVER_STAT_NAMES))
        _chk(lib().aifmm_ls_server_stats(self.s, _dp(out), len(out)))
        return dict(zip(SERVER_STAT_NAMES, out.tolist()))

    def close(self):
        if getattr(self, "s", None):
            lib().aifmm_ls_server_destroy(self.s)
            self.s = None

    def __del__(self):
        self.close()
