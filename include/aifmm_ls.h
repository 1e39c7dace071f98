/*
 * C ABI of the MI355X (gfx950) 2D Lippmann-Schwinger direct solver.
 *
 * Drop-in boundary for the reference's solve path (vaishna77/AIFMM_LS):
 * every entry point names the reference interface it replaces.  Plain
 * pointers and sizes only; complex arrays are interleaved (re, im) doubles,
 * matrices column-major.  Every function returns 0 on success and a
 * negative code on failure; aifmm_ls_last_error() then describes it.
 */
#ifndef AIFMM_LS_H
#define AIFMM_LS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aifmm_ls_problem aifmm_ls_problem;

typedef struct {
  int n_cheb;          /* Chebyshev nodes per leaf side (argv[1]) */
  int degree;          /* basis degree p (argv[3]) */
  int n_levels;        /* uniform tree depth (argv[7]) */
  int tol_pow;         /* ACA / RRQR tolerance 10^-tol_pow (argv[6]) */
  int qchoice;         /* contrast: 0..5 as LS_functions.hpp, 6 = bump sum (argv[5]) */
  int quantize_tables; /* 1: tables rounded to 6 significant digits like the text cache */
  double kappa;        /* wavenumber (argv[2]) */
  double L;            /* domain semi-length (argv[4]) */
  const double* bumps; /* qchoice 6: n_bumps x (a, w, cx, cy) */
  int n_bumps;
} aifmm_ls_params;

/* userkernel ctor + LS2DTree ctor (kernel.hpp:106-128, LS_Assembly.hpp:234-277):
 * points, contrast, far/near tables by device quadrature. */
int aifmm_ls_create(const aifmm_ls_params* p, aifmm_ls_problem** out);
/* The same with the tables read from the precomputation text cache under dir
 * (dir/precomputations/{M,Neighbor}/...; readMFromFile + readNeighborInteractions,
 * LS_Assembly.hpp:1217-1228, 1422-1459): the reference's ctor with its three
 * precomputation lines commented out (LS_Assembly.hpp:269-272, LS/Readme.md:1-5).
 * Every file a tree of n_levels levels reads must exist and have its shape; a
 * cache whose aifmm_key file names another kappa, L, n_cheb or degree is refused. */
int aifmm_ls_create_cached(const aifmm_ls_params* p, const char* dir, aifmm_ls_problem** out);
/* evaluatePrecomputations + writeMToFile + getNeighborInteractions
 * (LS_Assembly.hpp:269-272, 1166-1184, 1320-1420): far M at levels 0..10 and the
 * 53 neighbour operators (9 colleague, 12 fine, 20 separated fine, 12 coarse)
 * at levels 2..n_levels by device quadrature, written as the reference's text
 * cache under dir; *seconds (may be NULL) = time of the quadrature batch. */
int aifmm_ls_write_precomputations(aifmm_ls_problem* h, const char* dir, double* seconds);
/* The LS2DTree ctor with createAdaptiveTree + makeLevelRestriction_Tree in place of
 * createUniformTree (LS_Assembly.hpp:255-275, 609-742, 892-1072): a box splits
 * until the contrast interpolates from its Chebyshev nodes to its children's
 * to 10^-tree_adaptivity (the reference driver passes 1) and kappa r / (2 pi)
 * <= 1; then leaves two or more levels apart that are near are refined (2:1
 * restriction).  p->n_levels is unused.  Points, entries (getMatrixEntry with
 * the colleague, fine, separated-fine and coarse tables, LS_Assembly.hpp:474-541),
 * the right-hand side, the dense product and the fields work on it; assembly,
 * factorisation and the solves refuse it, since the reference's FMM2DTree
 * solver works on the uniform tree.  The reference's own tables stop at level 2
 * (it would read outside them for a leaf at level 1); here every leaf level's
 * tables are formed. */
int aifmm_ls_create_adaptive(const aifmm_ls_params* p, int tree_adaptivity, aifmm_ls_problem** out);
/* The adaptive tree's leaves in the reference's leafNodes order (leaf k holds
 * points 36k .. 36k+35): level, box number (children 4b + c), centre (x, y).
 * Any output may be null; *n = capacity in, the leaf count out. */
int aifmm_ls_leaves(aifmm_ls_problem* h, int* levels, int64_t* boxes, double* centres, int* n);
/* One precomputation table, unrounded (FarFieldInteraction::evaluateM for
 * family -1, LS2DTree::getOperator for families 0..3 = colleague, fine,
 * separated fine, coarse; FarFieldInteraction.hpp:153-160, LS_Assembly.hpp:1290-1317):
 * family -1: M at level (121 x R); else operator n of a target box at level (R x R). */
int aifmm_ls_table(aifmm_ls_problem* h, int family, int level, int n, double* out);
void aifmm_ls_destroy(aifmm_ls_problem* h);
const char* aifmm_ls_last_error(void);

/* userkernel::N (kernel.hpp:103) */
int64_t aifmm_ls_size(const aifmm_ls_problem* h);
/* userkernel::gridPoints (N x 2 interleaved) and contrast q(x_i) (LS_functions.hpp:99-112) */
int aifmm_ls_points(aifmm_ls_problem* h, double* xy, double* q);
/* precomputed tables: M (121 x R) and colleague operators T (9 x R x R); quad_capped
 * counts integrals that hit the refinement depth cap */
int aifmm_ls_tables(aifmm_ls_problem* h, double* M, double* T, int* quad_capped);
/* userkernel::getMatrixEntry batch (kernel.hpp:130-136, LS_Assembly.hpp:474-541) */
int aifmm_ls_entries(aifmm_ls_problem* h, int64_t n, const int64_t* I, const int64_t* J, double* out);
/* userKernel::RHSFunction with incidence direction (cos th, sin th) (LS_functions.hpp:113-120) */
int aifmm_ls_rhs_planewave(aifmm_ls_problem* h, double theta, double* b);

/* TOL_POW of the FMM2DTree ctor (FMM2DTree_find_error.hpp:98-112): the ACA / RRQR
 * tolerance 10^-tol_pow, settable until assemble (create takes an initial value). */
int aifmm_ls_set_tolerance(aifmm_ls_problem* h, int tol_pow);
/* FMM2DTree::getNodes + assemble_M2L + initialise_P2P_Leaf_Level
 * (FMM2DTree_find_error.hpp:1137-1361, 1729-1741) */
int aifmm_ls_assemble(aifmm_ls_problem* h);
/* skeleton ranks per box at a level after NCA truncation */
int aifmm_ls_ranks(aifmm_ls_problem* h, int level, int* ranks);
/* skeleton point lists (which: 0 in_chk, 1 in_chg, 2 out_chk, 3 out_chg); *n in: capacity, out: length */
int aifmm_ls_skeleton(aifmm_ls_problem* h, int level, int box, int which, int64_t* out, int* n);

/* FMM2DTree::eliminate_phase_efficient (FMM2DTree_find_error.hpp:2407-2451), here the
 * level-synchronous colour schedule.  precision: 0 = complex128, 1 = complex64. */
int aifmm_ls_factor(aifmm_ls_problem* h, int precision);
/* The same for n independent problems together: FMM2DTree::assemble_M2L ..
 * eliminate_phase_efficient (FMM2DTree_find_error.hpp:1338-1361, 2407-2451) of
 * every problem, where the reference factorises one problem per process.
 * Assembly runs concurrently per problem; the eliminations of all problems
 * then go out as one level-synchronous launch sequence (problems of different
 * depth start together at their leaf levels).  Each factorisation equals what
 * aifmm_ls_factor gives.  info (optional, 6 doubles): wall seconds in all,
 * wall seconds of the factorisations, launch rounds, kernel launches, seconds
 * spent launching, seconds between rounds (the members' host work).  On
 * failure the members that failed are left unfactored. */
int aifmm_ls_factor_batch(aifmm_ls_problem* const* hs, int n, int precision, double* info);
/* Zero-pivot epsilons of the next complex128 factorisation, in Eigen's
 * ColPivHouseholderQR "nonzero pivot" test (a pivot counts as zero below
 * eps x the largest column norm; the reference uses the FP64 epsilon,
 * FMM2DTree_find_error.hpp:2092 for L2M_f and :2477 for the base system):
 * l2mf_eps for the L2M_f pivot blocks (default 1e-12), base_eps for the
 * level-2 base system (default 0 = the FP64 epsilon).  Both in [0, 1). */
int aifmm_ls_set_pivot_eps(aifmm_ls_problem* h, double l2mf_eps, double base_eps);
/* Precision of the pivot blocks of the next complex64 factorisation, where the
 * reference factors them in FP64 (PartialPivLU / ColPivHouseholderQR of P2P[k][k]
 * and L2M_f, FMM2DTree_find_error.hpp:1819, 2092).  thr < 0: every
 * pivot block is widened and factored in FP64.  thr >= 0: each block of up to
 * 128 rows is first LU-factored in complex64 and keeps that factor when its
 * growth factor max|U|/max|A| times its condition estimate max|u_ii|/min|u_ii|
 * is at most thr; the rest are factored in FP64.  The choice is made on the
 * device per block.  Default 1e4: at N = 147,456 (kappa 40, 8 bumps) 34 of
 * 10,206 pivot blocks go to FP64, the refined solve keeps its 3 steps and
 * the factorisation takes 0.20 s instead of 0.21 s with every block in FP64. */
int aifmm_ls_set_growth_gate(aifmm_ls_problem* h, double thr);
/* Decisions of the last factorisation, for comparison with the reference's
 * (FMM2DTree_find_error.hpp:1424-1699 recompression ranks, :2092 pivots).
 * which 0: one record per recompressed box, 6 ints: level, colour step, box,
 * rank of [L2P | row fill-ins], rank of [P2M^H | column fill-ins^H], new rank.
 * which 1 (complex128): one record per L2M_f factorisation with zeroed
 * pivots, 4 ints: level, box, nonzero pivots, size.  *n gets the number of
 * ints; up to cap are written to out (out may be null). */
int aifmm_ls_factor_trace(aifmm_ls_problem* h, int which, int* out, int cap, int* n);
/* forward sweep + base solve + back_substitution_phase (FMM2DTree_find_error.hpp:2487-2532)
 * for nrhs right-hand sides (N x nrhs).  stats (>= 8 doubles, optional): solve seconds,
 * refinement steps, final relative residual max_c |b_c - A_H x_c| / |b_c|, ...
 * (with flexible GMRES converged and tol >= 1e-11, the recurrence's value of
 * that residual, which agrees with the FP64 product to rounding; otherwise
 * measured). */
int aifmm_ls_solve(aifmm_ls_problem* h, int nrhs, const double* b, double* x, double tol, int max_refine,
                   double* stats);
/* device-resident variant for timing: stage RHS, solve in HBM, fetch solution */
int aifmm_ls_set_rhs(aifmm_ls_problem* h, int nrhs, const double* b);
int aifmm_ls_solve_resident(aifmm_ls_problem* h, double tol, int max_refine, double* stats);
int aifmm_ls_get_solution(aifmm_ls_problem* h, double* x);
/* A batch of independent problems solved together: eliminate_phase_efficient's RHS
 * part + back_substitution_phase (FMM2DTree_find_error.hpp:2407-2532) of every
 * problem, with stage i of each problem's sweep merged into one launch (one
 * level-synchronous launch sequence for the batch), and likewise the FP64
 * matvecs of a flexible GMRES that runs every problem's Arnoldi process in
 * lockstep until each relative residual is <= tol (at most max_refine
 * preconditioner applications after x0 = M b).  One right-hand side
 * per problem: b[i], x[i] hold N_i complex values; relres (n, optional) gets
 * each problem's final relative residual; stats = time, steps, max relres.  All
 * problems must be factorised in the same precision. */
int aifmm_ls_solve_batch(aifmm_ls_problem* const* hs, int n, const double* const* b, double* const* x, double tol,
                         int max_refine, double* relres, double* stats);
/* Refinement method for max_refine > 0 (no reference counterpart: the reference
 * solves once, FMM2DTree_find_error.hpp:2407-2532): 0 = flexible GMRES with the
 * factorisation as right preconditioner (default), 1 = plain iterative
 * refinement x += M (b - A_H x).  Both use the FP64 operator of aifmm_ls_matvec. */
int aifmm_ls_set_refinement(aifmm_ls_problem* h, int method);
/* Dense FP64 product y = A x with the discretised operator itself, every
 * entry as getMatrixEntry over all leaf pairs: the generator's rhs = A x
 * (FMM2DTree_gen_rhs_x.hpp:2808-2903, getMatrix blocks at :2864-2871).  Far
 * entries are summed per source leaf through its moments (one Hankel series
 * per point and leaf); near entries are formed and zeroed below 1e-16 one by
 * one, as getMatrixEntry; unzeroed far entries below 1e-16 change a row sum
 * by less than 1e-16 |x_j| each.  Host buffers, N x nrhs column-major. */
int aifmm_ls_dense_matvec(aifmm_ls_problem* h, int nrhs, const double* x, double* y);
/* LS2DTree::findPhi (LS_Assembly.hpp:195, 507-512, 530-535) for
 * aifmm_ls_entries and the dense products: 1 (default), entries of the LS
 * operator I + kappa^2 q G; 0, of the volume potential G alone, which maps the
 * solution phi to the scattered field (hybrid.cpp:264-281 sets it to false for
 * that product).  Assembly and factorisation always use the LS operator. */
int aifmm_ls_set_find_phi(aifmm_ls_problem* h, int on);
/* userKernel::IncidenceFunction at the points (LS_functions.hpp:113-116):
 * u_i = exp(i kappa d.x_i), d = (cos theta, sin theta); N interleaved complex. */
int aifmm_ls_incident_field(aifmm_ls_problem* h, double theta, double* u);
/* The total field from a solution phi of the plane wave at theta
 * (hybrid.cpp:264-281): u = u_inc + G phi, with the dense volume-potential
 * product (findPhi false, every entry as getMatrixEntry); host buffers, N
 * interleaved complex each. */
int aifmm_ls_total_field(aifmm_ls_problem* h, double theta, const double* phi, double* u);
/* The dense product of aifmm_ls_dense_matvec without device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  x and y are the caller's device buffers (N x nrhs, column-major
 * interleaved complex); the scratch is a static 64 MiB device array of the
 * library (held on each GPU it runs on until the process ends), so N must
 * satisfy 121 x 4^L + N <= 4,194,304 complex values (up to 7 levels at 6 x 6
 * points per leaf); larger problems are refused.  stream: a hipStream_t to
 * launch on, returning at once, or null for the problem's own stream, waited
 * for before returning.  Calls that overlap on one GPU share the array and
 * race; consecutive calls need no clearing. */
int aifmm_ls_dense_matvec_noalloc(aifmm_ls_problem* h, int nrhs, const double* x, double* y, void* stream);
/* FP64 product y = A_H x with the assembled hierarchical operator (leaf P2P plus the
 * nested L2P M2L P2M far field), the operator the factorisation inverts; N x nrhs. */
int aifmm_ls_matvec(aifmm_ls_problem* h, int nrhs, const double* x, double* y);
/* statistics: out[0..] = t_tables, t_assemble, t_factor, factor flops, factor bytes,
 * base size, max rank, solve bytes (nrhs=1), recompressed boxes, operator bytes,
 * last solve time, NCA time, boxes whose sweeps keep the reference's order
 * (badly conditioned L2M_f), and the latest launch tuning of the active
 * factorisation's sweep (see aifmm_ls_set_tuning): run time with the tuned
 * workgroup sizes over the default's (1 when the default was kept; 0 before
 * the first solve), stages whose size was changed; then the top of the
 * one-column sweep applied as one dense matrix: its tree level (0 = every
 * level staged) and its size; the base factor's smallest over largest
 * |diagonal|; the number of L2M_f blocks with zeroed pivots (complex128);
 * complex64 with a growth gate: pivot blocks that kept their complex64 LU and
 * pivot blocks factored in FP64. */
int aifmm_ls_stats(aifmm_ls_problem* h, double* out, int n);
/* Launch tuning (no reference counterpart): the first solve at each RHS count
 * times the sweep's and the FP64 operator's stages at 256, 512 and 1024
 * threads per workgroup (with several right-hand sides also at 1, 2, 4 or 8
 * columns per pass), over whole runs of about 20 ms or less, and keeps
 * per-stage shapes that beat the default by 1 % over whole runs.  Results
 * then differ from untuned runs by summation order only.  1 = on (the
 * default, unless AIFMM_TUNE=0 is set), 0 = off, or a fixed shape for every
 * stage it fits, for tests of each shape: threads (256, 512, 1024) + 4096 x
 * columns per pass (0 = as many of 1, 2, 4 as fit; 8 at up to 512 threads).
 * Programs already tuned keep their shapes.  A new factorisation tunes
 * afresh. */
int aifmm_ls_set_tuning(int mode);
/* Device memory the handle holds (arena chunks: operator, factors, sweep and
 * refinement vectors).  No reference counterpart: the reference runs one
 * problem per process; a serving loop admits problems by this figure. */
int aifmm_ls_device_bytes(aifmm_ls_problem* h, int64_t* bytes);

/* LowRank::ACA_only_nodes (ACA.hpp:368-577) on explicit point lists: pivot rows rb
 * and columns cb (rank entries each), un-updated pivot columns Ac (n1 x rank) and rows
 * Ar (rank x n2), column-major interleaved complex. */
int aifmm_ls_aca(aifmm_ls_problem* h, int tol_pow, int n1, const int64_t* rows, int n2, const int64_t* cols,
                 int* rb, int* cb, double* Ac, double* Ar, int* rank);

/* Kernel self-test (no reference counterpart; for the test suite): runs one
 * batched dense kernel on host data.  op 0: C = A B (m x k, k x n); op 1:
 * C = A^{-1} B by LU (A m x m, B m x n); op 2: column-pivoted QR of A (m x n)
 * with relative threshold thr, iout[0] = rank, C = first k columns of Q; op 3:
 * C = A^{-1} B by column-pivoted QR (A m x m), iout[0] = nonzero pivots.
 * op + 10 runs the complex64 kernels.  op 14 (complex64 only): C = A^{-1} B
 * through the pivot-block growth gate at threshold thr (see
 * aifmm_ls_set_growth_gate), C[m n] (real part) = the block's growth-and-
 * condition estimate, iout[0] = 1 when the block went to FP64.  op 5 (op 15
 * in complex64): column-pivoted QR of A (m x n) at threshold thr by the
 * global-memory kernel (k & 1 = 0) or the register kernel (k & 1 = 1, at most
 * 36 rows and 256 columns), early stop when k & 2: C = the overwritten A (R
 * and the reflectors) followed by tau[min(m, n)], iout = rank, nonzero
 * pivots, then the column permutation (n entries). */
int aifmm_ls_selftest(int op, int m, int n, int k, const double* A, const double* B, double* C, int* iout,
                      double thr);
/* Box-program self-test (no reference counterpart; for the test suite): one
 * stage of nprog programs of m rows, y_p = S_p + sum_t alpha_t A_{p,t} X_{p,t}
 * (alpha_t = 1 for even t, -0.5 + 0.25i for odd t; saved to Y) and out_p =
 * B_p y_p, run in launch shape `shape` (see aifmm_ls_set_tuning; 0 = default)
 * in complex128 (precision 0) or complex64 (1).  Column-major interleaved
 * complex: A nprog*nterms blocks m x cols, X nprog*nterms blocks cols x nrhs,
 * S, Y, Out nprog blocks m x nrhs, Bm nprog blocks m x m. */
int aifmm_ls_selftest_prog(int precision, int shape, int nprog, int m, int nterms, int cols, int nrhs,
                           const double* A, const double* X, const double* S, const double* Bm, double* Y,
                           double* Out);

/* Stream self-test (no reference counterpart; for calibrating hardware traffic
 * counters): copies `bytes` between two fresh device buffers `reps` times in
 * alternating directions with the plain copy kernel k_stream_copy; *ms gets
 * the mean kernel time. */
int aifmm_ls_selftest_stream(int64_t bytes, int reps, double* ms);
/* Kernel timing (no reference counterpart; for benchmarks): when enabled, every
 * batched launch records a HIP event pair on its stream together with its
 * algorithmic bytes and flops.  profile_read fills, for up to max kernels,
 * out[5*i ..] = launches, summed milliseconds, bytes, flops, 0, and writes the
 * kernel names '\n'-separated into names (capacity names_len); returns the number
 * of kernels, or a negative code.  reset != 0 clears the totals afterwards. */
int aifmm_ls_profile_enable(int on);
int aifmm_ls_profile_read(int reset, int max, double* out, char* names, int names_len);
/* With profile_enable(2), each launch is also kept as a record: profile_launches
 * fills, for up to max launches in launch order, out[2*i ..] = algorithmic
 * bytes, milliseconds, and writes the kernel names '\n'-separated into names;
 * returns the number of records, or a negative code.  reset != 0 clears them. */
int aifmm_ls_profile_launches(int reset, int max, double* out, char* names, int names_len);

/* ---- Serving loop.  No reference counterpart: the reference factorises and
 * solves one problem per process (testFMM2D_find_error.cpp:41-176).  A server
 * keeps factorised problems resident on the GPU and runs requests on `slots`
 * worker threads, each on its problem's own HIP stream, so requests for
 * different problems overlap on the device; queued solves for one problem
 * with equal tol / max_refine are merged into one multi-RHS solve (at most 64
 * right-hand sides), and queued one-column solves of different problems with
 * equal tol / max_refine and factor precision into one batch solve (as
 * aifmm_ls_solve_batch, at most 32 problems).  New factorisations are admitted by device memory against
 * budget_bytes, evicting idle problems least-recently used first; a solve for
 * an evicted problem factorises it again.  Submission copies the inputs and
 * returns at once; x must stay valid until the request is waited for.  While a
 * server exists, sweeps launch directly instead of through HIP graphs. */
typedef struct aifmm_ls_server aifmm_ls_server;
int aifmm_ls_server_create(int64_t budget_bytes, int slots, aifmm_ls_server** out);
/* queue a factorisation (precision 0: complex128, 1: complex64); *problem names
 * the problem in later solves; *ticket (optional) is the request's ticket */
int aifmm_ls_server_factor(aifmm_ls_server* s, const aifmm_ls_params* p, int precision, int64_t* problem,
                           int64_t* ticket);
/* queue a solve of nrhs right-hand sides (N x nrhs interleaved complex, column-major) */
int aifmm_ls_server_solve(aifmm_ls_server* s, int64_t problem, int nrhs, const double* b, double* x, double tol,
                          int max_refine, int64_t* ticket);
/* block until the request ends; info (optional, 9 doubles) = status (0 ok,
 * -1 failed), latency (submit to end, s), queue wait (s), run time (s),
 * refinement steps, relative residual, right-hand sides in the merged solve,
 * 1 if the problem was factorised again for this request, problems in the
 * batch it ran in (1 unless merged with other problems' one-column solves).
 * Each ticket is waited for once. */
int aifmm_ls_server_wait(aifmm_ls_server* s, int64_t ticket, double* info);
/* drop a problem: its memory is freed now, or when its running request ends */
int aifmm_ls_server_release(aifmm_ls_server* s, int64_t problem);
/* out[0..10] = resident problems, resident bytes, budget, evictions,
 * re-factorisations, completed requests, executed batches, requests merged
 * into another's solve, most batches running at once, queued requests,
 * batches of several problems */
int aifmm_ls_server_stats(aifmm_ls_server* s, double* out, int n);
/* finish every queued request, then stop the slots and free all problems */
int aifmm_ls_server_destroy(aifmm_ls_server* s);

#ifdef __cplusplus
}
#endif
#endif
