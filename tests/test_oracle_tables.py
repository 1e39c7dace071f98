"""Tables and matrix entries of the oracle against direct quadrature.

Checks, with the 6-digit rounding switched off:
  * far table M = C Q^+ against a non-adaptive high-order tensor rule;
  * colleague operators T_n against direct quadrature (polar coordinates
    around the singular point for the self box);
  * far entries against Graf's addition theorem: the Hankel series equals the
    direct integral (i/4) kappa^2 q(x_i) int H_0(kappa|x_i - y|) phi_j(y) dy;
  * near entries use the table of the right neighbour slot;
and with rounding on, that every table value is its own 6-significant-digit
rounding.
"""
import numpy as np
import pytest
import scipy.special as sp

N_CHEB, DEG = 6, 6


def basis():
    return [(a, d - a) for d in range(DEG) for a in range(d + 1)]


def cheb_nodes():
    return -np.cos((np.arange(N_CHEB) + 0.5) / N_CHEB * np.pi)


def qpinv():
    xi = cheb_nodes()
    pts = np.array([(xi[i % N_CHEB], xi[i // N_CHEB]) for i in range(N_CHEB ** 2)])
    Q = np.array([[np.cos(a * np.arccos(x)) * np.cos(b * np.arccos(y)) for (a, b) in basis()] for x, y in pts])
    return np.linalg.pinv(Q)


def tensor_rule(n, panels):
    t, w = np.polynomial.legendre.leggauss(n)
    edges = np.linspace(-1, 1, panels + 1)
    xs = np.concatenate([0.5 * (edges[k + 1] - edges[k]) * t + 0.5 * (edges[k + 1] + edges[k]) for k in range(panels)])
    ws = np.concatenate([0.5 * (edges[k + 1] - edges[k]) * w for k in range(panels)])
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    W = np.outer(ws, ws)
    return X.ravel(), Y.ravel(), W.ravel()


def cheb_basis_vals(X, Y):
    return np.array([np.cos(a * np.arccos(X)) * np.cos(b * np.arccos(Y)) for (a, b) in basis()])


@pytest.fixture(scope="module")
def prob(oracle):
    return oracle.Problem(n_levels=3, kappa=10.0, tol_pow=8, qchoice=0, quantize=False)


def test_far_table_against_tensor_rule(prob):
    M, _, _ = prob.tables()
    r = 1.0 / 8
    kr = 10.0 * r
    X, Y, W = tensor_rule(20, 4)
    rho, th = np.hypot(X, Y), np.arctan2(Y, X)
    B = cheb_basis_vals(X, Y)
    P = qpinv()
    for n in (-60, -13, -2, 0, 1, 7, 31, 60):
        Cn = r * r * np.array([np.sum(W * sp.jv(n, kr * rho) * np.exp(-1j * n * th) * B[l]) for l in range(len(basis()))])
        ref = Cn @ P
        err = np.max(np.abs(M[n + 60] - ref))
        assert err < 2e-9 + 1e-6 * np.max(np.abs(ref)), (n, err)


def near_direct(kr, X0, Y0):
    """int_{[-1,1]^2} H0(kr |X - s|) T_l(s) ds for all bases l (polar split at X)."""
    t, w = np.polynomial.legendre.leggauss(40)
    out = np.zeros(len(basis()), complex)
    # four triangles with apex at X, one per square side
    corners = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
    for k in range(4):
        ax, ay = corners[k]
        bx, by = corners[(k + 1) % 4]
        # parametrise triangle (X, A, B): s = X + u((A + v(B - A)) - X), u, v in [0, 1]
        # graded radial variable u = s^4 removes the rho*log(rho) kink
        sg = 0.5 * (t + 1)
        u = sg ** 4
        wu = 0.5 * w * 4 * sg ** 3
        v = 0.5 * (t + 1)
        U, V = np.meshgrid(u, v, indexing="ij")
        WU, WV = np.meshgrid(wu, 0.5 * w, indexing="ij")
        ex = ax + V * (bx - ax) - X0
        ey = ay + V * (by - ay) - Y0
        SX, SY = X0 + U * ex, Y0 + U * ey
        jac = np.abs(ex * (by - ay) - ey * (bx - ax)) * U
        dist = np.hypot(SX - X0, SY - Y0)
        H = sp.hankel1(0, kr * dist)
        Bv = cheb_basis_vals(np.clip(SX, -1, 1), np.clip(SY, -1, 1))
        out += np.array([np.sum(WU * WV * jac * H * Bv[l]) for l in range(len(basis()))])
    return out


NC = [(-2, -2), (0, -2), (2, -2), (2, 0), (2, 2), (0, 2), (-2, 2), (-2, 0), (0, 0)]


@pytest.mark.parametrize("slot", [0, 3, 8])
def test_near_table_against_direct(prob, slot):
    _, T, _ = prob.tables()
    r = 1.0 / 8
    kr = 10.0 * r
    xi = cheb_nodes()
    P = qpinv()
    for p in (0, 7, 21, 35):
        X0 = xi[p % N_CHEB] - NC[slot][0]
        Y0 = xi[p // N_CHEB] - NC[slot][1]
        if slot == 8:
            V = near_direct(kr, X0, Y0)
        else:
            Xg, Yg, W = tensor_rule(20, 4)
            B = cheb_basis_vals(Xg, Yg)
            H = sp.hankel1(0, kr * np.hypot(Xg - X0, Yg - Y0))
            V = np.array([np.sum(W * H * B[l]) for l in range(len(basis()))])
        ref = (1j * r * r / 4) * V @ P
        err = np.max(np.abs(T[slot][p] - ref))
        assert err < 5e-9 + 2e-6 * np.max(np.abs(ref)), (slot, p, err)


def test_entries_graf_identity_and_slots(prob):
    """Far entries (Hankel series) and near entries (tables) both equal the
    same direct integral (i/4) kappa^2 q_i r^2 sum_l V_l(X) Q^+(l, j%R)."""
    xy, q = prob.points()
    R = 36
    r = 1.0 / 8
    kr = 10.0 * r
    P = qpinv()
    rng = np.random.default_rng(7)
    # points near the contrast peak so that entries are not ~0
    cand = np.where(q > 1e-3)[0]
    I = rng.choice(cand, 12)
    J = rng.integers(0, prob.N, 12)
    J = np.concatenate([J, I + 1 - 2 * (I % 2)])  # include near pairs
    I = np.concatenate([I, I])
    vals = prob.entries(I, J)
    Xg, Yg, W = tensor_rule(24, 6)
    B = cheb_basis_vals(Xg, Yg)
    for i, j, v in zip(I, J, vals):
        lj = j // R
        cx = np.mean(xy[lj * R:(lj + 1) * R, 0])
        cy = np.mean(xy[lj * R:(lj + 1) * R, 1])
        X0, Y0 = (xy[i, 0] - cx) / r, (xy[i, 1] - cy) / r
        if max(abs(X0), abs(Y0)) < 1.5:
            V = near_direct(kr, X0, Y0)
        elif max(abs(X0), abs(Y0)) < 3.2:
            V = np.zeros(len(basis()), complex)
            for qx in (-1, 1):
                for qy in (-1, 1):  # quarter boxes for accuracy near the source
                    V += 0.25 * np.array([np.sum(W * sp.hankel1(0, kr * np.hypot(0.5 * Xg + 0.5 * qx - X0, 0.5 * Yg + 0.5 * qy - Y0))
                                                 * cheb_basis_vals(0.5 * Xg + 0.5 * qx, 0.5 * Yg + 0.5 * qy)[l]) for l in range(len(basis()))])
        else:
            H = sp.hankel1(0, kr * np.hypot(Xg - X0, Yg - Y0))
            V = np.array([np.sum(W * H * B[l]) for l in range(len(basis()))])
        ref = (1j / 4) * 10.0 ** 2 * q[i] * r * r * (V @ P)[j % R] + (1.0 if i == j else 0.0)
        # the reference accepts each adaptive panel at 1e-7 absolute; around the
        # log singularity of the self box the accepted panels of ~15 refinement
        # levels each contribute, so allow 100 eps there (10 eps elsewhere),
        # propagated through (i/4) kappa^2 q r^2 sum_l |Q^+(l, j%R)|
        nacc = 100 if i // R == j // R else 10
        bound = nacc * 1e-7 * 0.25 * 10.0 ** 2 * abs(q[i]) * r * r * np.sum(np.abs(P[:, j % R])) + 1e-12
        assert abs(v - ref) < bound, (i, j, v, ref, bound)


def test_quantised_tables_are_six_digit(oracle):
    p = oracle.Problem(n_levels=2, kappa=10.0, tol_pow=8, qchoice=0, quantize=True)
    M, T, _ = p.tables()
    for arr in (M.ravel(), T.ravel()):
        for z in arr[::97]:
            for v in (z.real, z.imag):
                assert float("%.6g" % v) == v
