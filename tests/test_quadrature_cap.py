"""The leaf tables' adaptive quadrature (k_tables.hip) stops refining a panel
only at its tolerance, unless the panel tree reaches its depth or stack cap;
a capped integral is counted (aifmm_ls_tables' quad_capped).  The reference's
quadrature has no cap (domain2D.hpp:69-120), so at every benchmark
configuration no integral may hit it."""
import numpy as np
import pytest

import aifmm_ls_amd as A
from aifmm_ls_amd.workloads import gaussian_bumps

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(1)
CONFIGS = [  # name, levels, kappa, tolerance power, contrast choice, bumps
    ("C1 reference run 6 10 6 1 0 8 5", 5, 10.0, 8, 0, None),
    ("C2 N=147,456 kappa 40", 6, 40.0, 7, 6, gaussian_bumps(2024, 8)),
    ("C3 smallest and largest", 3, 5.0, 6, 6, gaussian_bumps(1, 3)),
    ("C3 largest kappa", 5, 40.0, 6, 6, gaussian_bumps(2, 6)),
    ("C4 N=589,824 kappa 60", 7, 60.0, 6, 6,
     np.c_[RNG.uniform(0.2, 1.0, 4), RNG.uniform(0.15, 0.25, 4), RNG.uniform(-0.5, 0.5, (4, 2))]),
    ("C5 largest kappa", 6, 60.0, 7, 6, gaussian_bumps(3, 6)),
]


@pytest.mark.parametrize("cfg", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_no_capped_integral(cfg):
    name, L, kappa, tp, qc, bumps = cfg
    g = A.LSProblem(L, kappa, tp, qc, bumps=bumps)
    M, T, capped = g.tables()
    print(f"  {name}: N = {g.N}, capped integrals {capped}", flush=True)
    assert capped == 0
    assert np.all(np.isfinite(M)) and np.all(np.isfinite(T))
    g.close()
