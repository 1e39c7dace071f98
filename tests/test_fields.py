"""Scattered and total fields (the reference's findPhi = false operator,
LS_Assembly.hpp:507-512, 530-535, and hybrid.cpp:264-281): the volume
potential G's entries and dense product against the oracle's, and the
identity that ties G to the LS operator A = I + kappa^2 diag(q) G: for the
exact solution phi of A phi = -kappa^2 q u_inc, phi = -kappa^2 q u_tot with
u_tot = u_inc + G phi, point by point.  Bars: entries to rounding (1e-11
relative, 1e-13 absolute under 1e-14); products 1e-12; the identity to
1e-11 of |phi| (entries zeroed below 1e-16 enter A and G differently)."""
import numpy as np
import pytest

import aifmm_ls_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BUMPS = [[1.0, 0.15, 0.1, 0.2], [0.7, 0.1, -0.3, -0.2]]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("L,kap,qc,bumps", [(2, 10.0, 0, None), (3, 20.0, 6, BUMPS)])
def test_volume_potential_entries_and_product(L, kap, qc, bumps):
    print(f"\n[case] G at L={L} kappa={kap}", flush=True)
    g = A.LSProblem(L, kap, 8, qc, bumps=bumps)
    o = O.Problem(L, kap, 8, qc, bumps=bumps)
    g.set_find_phi(False)
    o.set_find_phi(False)
    rng = np.random.default_rng(L)
    I = rng.integers(0, g.N, 20000)
    J = rng.integers(0, g.N, 20000)
    eg, eo = g.entries(I, J), o.entries(I, J)
    big = np.abs(eo) > 1e-14
    assert np.max(np.abs(eg - eo)[big] / np.abs(eo)[big]) < 1e-11
    assert np.all(np.abs(eg - eo)[~big] < 1e-13)
    G = o.dense()
    x = rng.standard_normal(g.N) + 1j * rng.standard_normal(g.N)
    d = rel(g.dense_matvec(x), G @ x)
    print(f"  G x vs oracle {d:.2e}", flush=True)
    assert d < 1e-12
    g.set_find_phi(True)
    # back on the LS operator
    assert np.allclose(g.entries(I[:50], J[:50]), O.Problem(L, kap, 8, qc, bumps=bumps).entries(I[:50], J[:50]),
                       rtol=1e-11, atol=1e-13)
    g.close()


def test_total_field_identity():
    """3 levels, kappa 20, two bumps: phi = A^{-1} b exactly (dense LU of the
    oracle's matrix); the GPU's total field then gives -kappa^2 q u_tot = phi."""
    L, kap = 3, 20.0
    g = A.LSProblem(L, kap, 8, 6, bumps=BUMPS)
    o = O.Problem(L, kap, 8, 6, bumps=BUMPS)
    Ad = o.dense()
    b = o.rhs(0.7)
    phi = np.linalg.solve(Ad, b)
    u_inc = g.incident_field(0.7)
    assert rel(u_inc, o.incident(0.7)) < 1e-15
    u = g.total_field(phi, 0.7)
    _, q = g.points()
    lhs = -kap * kap * q * u
    d = np.max(np.abs(lhs - phi)) / np.max(np.abs(phi))
    print(f"\n  max |phi + kappa^2 q u_tot| / max |phi| = {d:.2e}", flush=True)
    assert d < 1e-11
    # the scattered field alone equals u_tot - u_inc
    assert rel(g.scattered_field(phi), u - u_inc) < 1e-14
    g.close()
