"""Named workloads for profiling runs: ``python -m aifmm_ls_amd.workloads NAME``.

Each one builds the 6-level problem of the mixed-precision solve benchmark
(N = 147,456 unknowns, kappa = 40, compression tolerance 1e-7, 8 Gaussian
bumps drawn from seed 2024) and runs one part of the solve path, so that a
profiler sees that part on its own:

  sweep   a complex64 factorisation, then three direct sweeps
          (``solve_resident(tol, 0)``: no refinement)
  factor  one assembly, then two complex64 and two complex128
          factorisations
  solve   a complex64 factorisation, the residual after 1..5 refinement
          steps, then three refined solves (tolerance 1e-7)
  stream  no problem: three copies of 1 GiB between two fresh 1 GiB device
          buffers with the plain copy kernel k_stream_copy (16-byte elements),
          for calibrating FETCH_SIZE / WRITE_SIZE against a stream of known
          size far beyond the L2 and MALL capacities

Run under ``rocprofv3``; with ``AIFMM_GRAPH=0`` the sweeps' launches are
issued one by one instead of replayed as HIP graphs."""
import argparse
import time
import sys

import numpy as np


def gaussian_bumps(seed, n):
    """n Gaussian contrast bumps (amplitude, width, centre x, centre y) drawn
    from `seed`: the contrast of the benchmark problems."""
    rng = np.random.default_rng(seed)
    return np.c_[rng.uniform(0.2, 1.5, n), rng.uniform(0.05, 0.2, n), rng.uniform(-0.7, 0.7, (n, 2))]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("name", choices=("sweep", "factor", "solve", "stream"))
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--kappa", type=float, default=40.0)
    ap.add_argument("--tol-pow", type=int, default=7)
    ap.add_argument("--bumps", type=int, default=8)
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args(argv)

    import aifmm_ls_amd as A

    if args.name == "stream":
        ms = A.selftest_stream(1 << 30, 3)
        print(f"stream: 1 GiB copy {ms:.3f} ms, {2 * (1 << 30) / ms / 1e6:.1f} GB/s read+write", flush=True)
        return 0
    g = A.LSProblem(args.levels, args.kappa, args.tol_pow, qchoice=6, bumps=gaussian_bumps(args.seed, args.bumps))
    t0 = time.perf_counter()
    g.assemble()
    print(f"assembly: {time.perf_counter() - t0:.3f} s", flush=True)
    if args.name == "factor":
        for prec, name in ((A.PREC_C64, "complex64"), (A.PREC_C128, "complex128")):
            for i in range(2):
                g.factor(prec)
                print(f"{name} factor {i}: {g.stats()['t_factor']:.3f} s", flush=True)
    elif args.name == "solve":
        g.factor(A.PREC_C64)
        g.set_rhs(g.rhs(0.3))
        for k in range(0, 6):
            st = g.solve_resident(1e-30, k)
            print(f"after {k} steps: relres {st[2]:.3e}", flush=True)
        for i in range(3):
            st = g.solve_resident(1e-7, 30)
            print(f"solve {i}: {1e3 * st[0]:.3f} ms, {st[1]:.0f} steps, relres {st[2]:.3e}", flush=True)
    else:
        g.factor(A.PREC_C64)
        g.set_rhs(g.rhs(0.3))
        for i in range(3):
            st = g.solve_resident(10.0 ** -args.tol_pow, 0)
            print(f"sweep {i}: {1e3 * st[0]:.3f} ms", flush=True)
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
