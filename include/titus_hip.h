/*
 * titus_hip.h -- C-ABI of the MI355X (gfx950) DLRA continuous-slowing-down
 * transport step.
 *
 * The reference (CSMMLab/TITUS, Julia) drives this path through a mutable
 * solver object:
 *
 *   SolverCSD(settings)                                 code/SolverCSD.jl:63-146
 *   SolveFirstCollisionSourceDLR(obj)                   code/SolverCSD.jl:891-1120
 *   SolveFirstCollisionSourceDLRBUGRejection(obj)       code/SolverCSD.jl:2011-2374
 *
 * Each entry point below replaces one piece of that interface.  All arrays
 * are FP64 and column-major (Julia layout), so a Julia `ccall` passes its
 * arrays without copies.  Spatial cells are numbered c = i*ny + j (0-based),
 * the reference's vectorIndex (code/utils.jl:98-100) minus one.
 *
 * Conventions: every function returns 0 on success or a negative TITUS_E_*
 * code; titus_last_error() then describes the failure (thread-local).  No
 * exception crosses the ABI.  A context owns one HIP stream and device
 * buffers; it is not shared between threads; separate contexts are
 * independent.  The library copies every input at titus_create; the caller
 * keeps ownership of its buffers.
 */
#ifndef TITUS_HIP_H
#define TITUS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TITUS_ABI_VERSION 3

typedef struct titus_ctx titus_ctx;

enum {
  TITUS_OK = 0,
  TITUS_E_ARG = -1,        /* invalid argument or inconsistent sizes */
  TITUS_E_HIP = -2,        /* a HIP runtime call failed */
  TITUS_E_NONFINITE = -3,  /* a non-finite value appeared in the state */
  TITUS_E_RANK = -4,       /* rank bookkeeping failed (cap, completion) */
  TITUS_E_STATE = -5,      /* call out of order (e.g. step past the grid) */
  TITUS_E_NODEVICE = -6    /* no usable GPU */
};

/* FIXED: SolveFirstCollisionSourceDLR (SolverCSD.jl:891-1120); BUG:
 * ...DLRBUGRejection (:2011-2374; rejection off unless titus_set_rejection
 * turns it on); PARALLEL:
 * SolveFirstCollisionSourceDLRParallel (:1726-2008). */
enum { TITUS_INTEGRATOR_FIXED = 0, TITUS_INTEGRATOR_BUG = 1, TITUS_INTEGRATOR_PARALLEL = 2,
       /* CudaSolveFirstCollisionSourceDLR4thOrder (:1422-1724): fixed rank,
        * central stencils only with RK4 K/L/S substeps, the analytic
        * uncollided flux of titus_set_analytic_beam; 2D */
       TITUS_INTEGRATOR_RK4 = 3 };
/* FP64: every operation in FP64.
 * MIXED: the uncollided flux psi stored and swept in FP32 (half the HBM
 *   traffic of the dominant kernel); everything else in FP64.
 * MIXED_FULL_FP32: MIXED plus the streaming stencil, the psi contractions and
 *   the Galerkin products X^T L X in FP32 with FP32 accumulation.  The
 *   rank-adaptive solve amplifies their rounding on heterogeneous phantoms
 *   (dose ~1e-3 from FP64), so it is not the default mixed form. */
enum { TITUS_PRECISION_FP64 = 0, TITUS_PRECISION_MIXED = 1, TITUS_PRECISION_MIXED_FULL_FP32 = 2 };

/* Problem description: the operands SolverCSD(settings) builds plus the
 * ordinate reduction and initial factors of the Solve* functions
 * (code/SolverCSD.jl:63-146, 904-961, 2025-2082). */
typedef struct titus_problem_v1 {
  int32_t version;            /* TITUS_ABI_VERSION (1..3 accepted: fields they lack default) */
  int32_t nx, ny;             /* NCellsX, NCellsY */
  double dx, dy;              /* settings.dx, settings.dy */
  const double* density;      /* C = nx*ny*nz, densityVec (SolverCSD.jl:143) */

  int32_t nm;                 /* moments carried */
  const int32_t* degree;      /* nm: Legendre degree l of each moment */
  const double* Ax;           /* nm x nm  pn.Ax    (PNSystem.jl:142-194) */
  const double* Az;           /* nm x nm  pn.Az */
  const double* absAx;        /* nm x nm  AbsAx    (SolverCSD.jl:83-90) */
  const double* absAz;        /* nm x nm  AbsAz    (SolverCSD.jl:92-97) */

  int32_t nq;                 /* reduced ordinates */
  const double* omega;        /* nq x 3   qReduced (SolverCSD.jl:922) */
  const double* M;            /* nm x nq  MReduced (SolverCSD.jl:923) */

  int32_t ne;                 /* energy grid points, csd.eTrafo */
  double dE;                  /* eTrafo[2]-eTrafo[1] (SolverCSD.jl:960) */
  const double* stopping;     /* ne, csd.S */
  int32_t nsigma;             /* nPN+1 */
  const double* sigma;        /* ne x nsigma, row-major: sigma_l(eGrid[n]) */

  int32_t nb;                 /* ghost cells */
  const int32_t* bcells;      /* nb distinct ghost cells (SolverCSD.jl:108-131) */
  const double* bc_base;      /* nb x nq: PsiBeam without its energy factor */
  const double* bc_scale;     /* ne: energy factor of PsiBeam at eGrid[n] */
  int32_t nrestore;           /* cells whose ordinate-0 flux survives the sweep */
  const int32_t* restore_cells;  /* (SolverCSD.jl:1002,1005) */
  int32_t ncand;              /* canonical-completion candidates (interior cells) */
  const int32_t* cand_cells;

  int32_t integrator;         /* TITUS_INTEGRATOR_* */
  int32_t precision;          /* TITUS_PRECISION_* */
  int32_t r;                  /* initial rank */
  int32_t rank_cap;           /* settings.r for the adaptive integrator */
  double theta, theta_index;  /* settings.ϑ, settings.ϑIndex */
  double rank_tol;            /* relative singular-value floor of the bases */
  const double* X0;           /* C x r  (NULL: svd!(zeros) start, ghost rows 0) */
  const double* S0;           /* r x r  (NULL: zeros) */
  const double* W0;           /* nm x r (NULL: identity columns) */

  /* version 2: a third spatial axis.  The reference is 2D; a 3D grid has
   * cells c = (i*ny + j)*nz + k with the k axis paired with Omega_y (Ay).
   * Read only when version >= 2; version-1 callers describe 2D grids. */
  int32_t nz;                 /* cells along k (1: 2D) */
  double dz;
  const double* Ay;           /* nm x nm  pn.Ay (PNSystem.jl:142-194), nz > 1 only */
  const double* absAy;        /* nm x nm  |Ay| (Roe, as SolverCSD.jl:83-97), nz > 1 only */

  /* version 3: the initial uncollided flux psi = SetupIC(obj, ...)
   * (SolverCSD.jl:155-185, 916), nonzero for the "waterBeam" problem. */
  const double* psi0;         /* C x nq (NULL: zeros) */
  /* version 3: uncollided flux scheme.  1 (or 0): first-order upwind,
   * solveFluxUpwind! (SolverCSD.jl:388-463); 2: the second-order limited
   * flux solveFlux! (:295-386) of CSolveFirstCollisionSourceDLR
   * (:1122-1360).  2D grids; the adjoint needs 1. */
  int32_t flux_order;
} titus_problem_v1;

/* SolverCSD(settings): validate, copy operands to device `device`.
 * Refused with TITUS_E_ARG (before any device call): negative counts, a
 * missing operand (restore_cells when nrestore > 0 included), a density that
 * is not positive and finite, a non-finite initial flux, cell indices out of
 * range, a non-positive energy step. */
int titus_create(struct titus_ctx** out, const titus_problem_v1* problem, int32_t device);

/* The explicit scheme's CFL number dE / min(dx, dy[, dz]) * max(1 / rho),
 * which the reference prints before its energy loop (SolverCSD.jl:963, 2084).
 * Above 1 the explicit upwind/Roe update can grow without bound; the library
 * still runs (as the reference does) and a caller decides whether to warn. */
int titus_get_cfl(const struct titus_ctx* ctx, double* cfl);

/* One iteration `for n = ...` of the energy loop (SolverCSD.jl:972-1114 /
 * 2102-2369).  Steps run in order 1..ne-1; `step` must be the next one. */
int titus_step(struct titus_ctx* ctx, int32_t step);

/* Run the remaining steps up to and including `last_step` (<= 0: to the end
 * of the grid as the reference's loop does). */
int titus_solve(struct titus_ctx* ctx, int32_t last_step);

/* Run `n` independent solves (e.g. pencil beamlets of a treatment plan) to
 * the end of their energy grids, `nthreads` host workers at a time (<= 0:
 * one per hardware thread, at most n).  Each context keeps its own stream, so
 * the beamlets' kernels and host algebra overlap.  status[i] (optional)
 * receives each solve's return code; the first failure is returned.  A
 * context may appear only once (TITUS_E_ARG otherwise, before any solve).
 * No reference counterpart: the reference solves one beam per call
 * (code/SolverCSD.jl:891, 2011). */
int titus_solve_batch(struct titus_ctx** ctxs, int32_t n, int32_t nthreads, int32_t* status);

/* The same batch with shared launches: up to `width` solves (<= 0: all n)
 * advance side by side, one host thread each, in `groups` groups (<= 0: 1)
 * of about width / groups solves with a stream each; in a group every kernel
 * of the energy step runs as one launch for all solves that reach it
 * together (solves of different ranks, ordinate counts and energy grids
 * included), and the groups' host work overlaps the other groups' device
 * work.  A finished solve's place goes to the next context.  Each solve's
 * arithmetic is that of titus_solve on its own (bitwise equal results).
 * Contexts must be 2D, on one device, with the fixed-rank, BUG or parallel
 * integrator, the upwind flux and no checkpoints (TITUS_E_ARG otherwise,
 * before any solve).  status[i] (optional) as titus_solve_batch; stats
 * (optional, 6 values): flushes (host syncs of the groups), batched
 * launches, the recorded launches they carried, nanoseconds spent in flushes
 * and, of those, waiting for the streams, and the sum over flushes of the
 * longest recorded list (the fewest launches those flushes could take).
 * Groups are buckets of similar contexts (moment count, rank capacity); each
 * tunes its tiles while it runs: over its first 40 flushes it times the
 * matrix-core GEMM and K-step at 2, 4 and 8 16-cell tiles per wavefront and
 * the sweep at strips of 2, 4, 8 rows (and non-temporal streaming), then
 * keeps the fastest of each.  These choices regroup the same tiles into
 * workgroups and change no result.  tiles (optional, 4 per group, groups
 * clamped to [1, width]): the chosen GEMM and K-step tiles per wavefront,
 * the sweep's rows per strip and non-temporal flag.
 * No reference counterpart: the reference solves one beam per call
 * (code/SolverCSD.jl:891, 2011). */
int titus_solve_batch_lockstep(struct titus_ctx** ctxs, int32_t n, int32_t width, int32_t groups,
                               int32_t* status, int64_t* stats, int32_t* tiles);

/* The pencil beam of the RK4 integrator's analytic uncollided flux
 * (SolverCSD.jl:1437-1440, 1549-1557):
 *   psi[c, q] = 1e5 exp(-1e4 (Omega1 - w_x)^2) exp(-1e4 (Omega3 - w_z)^2)
 *               exp(-int sigma) normpdf(x_i - tau w_x, x0, 0.1)
 *               normpdf(y_j - tau w_z, y0, 0.1)
 * at tau = eTrafo[n] with int sigma = dE sum_{k <= n} sigma_0(E_k), written
 * where both normpdf factors are >= 1e-6 (other entries keep their value). */
typedef struct titus_beam_v1 {
  double x0, y0;              /* settings.x0, settings.y0 (pos_beam) */
  double Omega1, Omega3;      /* settings.Omega1, settings.Omega3 */
  const double* xmid;         /* nx: settings.xMid */
  const double* ymid;         /* ny: settings.yMid */
  const double* tau;          /* ne: csd.eTrafo */
} titus_beam_v1;
/* Set the analytic beam of a TITUS_INTEGRATOR_RK4 context (before its first
 * step; required by it). */
int titus_set_analytic_beam(struct titus_ctx* ctx, const titus_beam_v1* beam);

/* Rejection of the rank-adaptive step (SolverCSD.jl:2204-2233, 2282-2301),
 * BUG integrator only, before the first step.  c_eta > 0 turns it on (the
 * reference's settings.cη, 5 by default); 0 turns it off (the default).  A
 * step is rejected when its truncation keeps all 2r columns or its eta
 * exceeds c_eta * theta * max(1e-7, |S|^theta_index) / dE, and 2r is below
 * the rank cap (the reference's `rmax` is undefined; settings.r is used).
 * Unlike the reference, which repeats the energy index with the dose and
 * the swept psi already applied, a rejected step restarts from its own start
 * (X, S, W, psi, dose) with the rejected half's new directions added to X
 * and W and zeros to S -- rank 2r after the streaming half, r0 + r after the
 * scattering half -- so nothing is counted twice.  The history then holds
 * the rank each accepted attempt started from. */
int titus_set_rejection(struct titus_ctx* ctx, double c_eta);
/* Number of rejected attempts so far. */
int titus_get_rejections(const struct titus_ctx* ctx, int64_t* count);

/* Number of steps taken so far and the current rank. */
int titus_progress(const struct titus_ctx* ctx, int32_t* steps_done, int32_t* rank);

/* Copy the state out (NULL pointers are skipped):
 *   dose C; X C x rank; S rank x rank; W nm x rank; psi C x nq (cell-major
 *   columns, Mat2Vec layout). */
int titus_get_state(struct titus_ctx* ctx, double* dose, double* X, double* S, double* W,
                    double* psi);

/* Rank history and the streaming eta indicator, n entries each at most.
 * BUG: the rank before each step (`rankInTime`, SolverCSD.jl:2109) and eta
 * (SolverCSD.jl:2220).  PARALLEL: the rank after each step (`rVec[2,:]`,
 * SolverCSD.jl:1946); no eta. */
int titus_get_history(const struct titus_ctx* ctx, int32_t* ranks, double* etas, int32_t n);

/* Device time of the library's kernels since timing was (re)enabled, measured
 * with HIP events on the context's stream: total milliseconds, launch count,
 * and the algorithmic HBM bytes and flops of those launches, for kernel
 * `name` ("bc", "dose", "sweep", "stream_K", "galerkin", "gram", "gemm",
 * "chol").  `period`: 0 off; n > 0 times the launches of every n-th energy
 * step (1: all of them; a larger period keeps the event overhead out of most
 * steps).  Host-side phases are reported under the names "host:step" (wall
 * time of the sampled steps), "host:wait" (blocked on the stream) and
 * "host:<phase>" for the host algebra, in wall milliseconds. */
int titus_set_timing(struct titus_ctx* ctx, int32_t period);
int titus_kernel_stats(const struct titus_ctx* ctx, const char* name, double* total_ms,
                       int64_t* launches, double* bytes, double* flops);

/* Launch plan the library uses for one of this context's kernels (no
 * counterpart in the reference: its loops have no launch parameters).
 * kernel "sweep": values[0] = rows per wavefront strip, values[1] = ordinate
 * groups, values[2] = 1 when the plan was chosen by timing the candidates at
 * titus_create (0: the size heuristic, for problems too small to tune),
 * values[3] = 1 when psi streams through non-temporal loads and stores.
 * Writes at most n values. */
int titus_get_tuning(const struct titus_ctx* ctx, const char* kernel, int32_t* values, int32_t n);

/* Running hash (FNV-1a) of the solve's discrete decisions -- kept rank and
 * completion vectors of every basis update, truncation ranks -- and their
 * count.  No reference counterpart.  Two solves with equal hashes took the
 * same decisions, so a central difference between them differentiates one
 * smooth branch of the forward map. */
int titus_get_decisions(const struct titus_ctx* ctx, uint64_t* hash, int64_t* count);

/* ---- adjoint (no reference counterpart: the reference has no gradients) ----
 * The differentiated map is this library's fixed-rank solve
 * (SolveFirstCollisionSourceDLR, code/SolverCSD.jl:891-1120): the dose after
 * the steps taken, as a function of the density (problem->density), with the
 * energy grid, beam source and P_N operators held fixed.  FP64, fixed-rank
 * integrator, upwind flux and 2D grids only (TITUS_E_ARG otherwise).
 *
 * titus_set_checkpoints: before the first step, keep the state (X, psi on the
 * device; S, W on the host) before steps 1, 1 + every, 1 + 2 every, ... of the
 * forward solve.  TITUS_E_STATE after a step has been taken. */
int titus_set_checkpoints(struct titus_ctx* ctx, int32_t every);

/* Gradient of <dose_bar, dose> with respect to the density (C values each):
 * d/d rho_c.  Recomputes each checkpointed segment once with its intermediates
 * kept, then runs it backwards; the forward state (dose, factors, psi) is
 * unchanged afterwards, also when the call fails (TITUS_E_NONFINITE when the
 * reverse pass overflows).  TITUS_E_STATE without checkpoints or steps;
 * TITUS_E_ARG for a non-finite dose_bar.  The rrule of a dose objective
 * J(dose) calls it with dose_bar = dJ/d dose. */
int titus_adjoint(struct titus_ctx* ctx, const double* dose_bar, double* grad_density);

/* titus_adjoint over n contexts on up to nthreads host threads (as
 * titus_solve_batch); dose_bars[i] and grads[i] belong to ctxs[i]. */
int titus_adjoint_batch(struct titus_ctx** ctxs, int32_t n, int32_t nthreads,
                        const double* const* dose_bars, double* const* grads, int32_t* status);

/* The same gradients with shared launches: each context's set-up (saved
 * forward state, seed) and wrap-up (restored state, gradient) run on its own
 * stream; the recomputation of its checkpoint segments and the reverse sweep
 * run in a lockstep batch (as titus_solve_batch_lockstep: `width` at once in
 * `groups` groups, every kernel one launch per group), bitwise the gradients
 * of titus_adjoint.  A forward solve with checkpoints may itself run in
 * titus_solve_batch_lockstep.  No reference counterpart (the reference has
 * no gradients). */
int titus_adjoint_batch_lockstep(struct titus_ctx** ctxs, int32_t n, int32_t width, int32_t groups,
                                 const double* const* dose_bars, double* const* grads,
                                 int32_t* status);

void titus_destroy(struct titus_ctx* ctx);

/* ---- full-rank P_N solver ---------------------------------------------------
 * SolveFirstCollisionSource(obj), code/SolverCSD.jl:465-586: the dense C x nm
 * moment matrix u advanced by explicit-Euler streaming and an implicit
 * per-cell scattering solve, beside the same uncollided sweep.  It differs
 * from the DLR path where the reference does: no 1/rho in the uncollided flux
 * (:551), the beam source at energy[n] (:530-537), the in-scattering source
 * inside the scattering solve (:564), and the dose
 * dE (u[:,1] + uOUnc) SMid[n-1] / rho / (1 + [n == 2 or n == nEnergies]) (:575).
 * FP64 and 2D.  Square grids only: the scattering loop numbers cells
 * (i-1)*nx + j (:563), the cell numbering only when nx == ny.
 * Conventions (status codes, titus_last_error, one stream per context) as
 * above. */
typedef struct titus_fcs titus_fcs;

typedef struct titus_fcs_problem_v1 {
  int32_t version;            /* TITUS_ABI_VERSION */
  int32_t nx, ny;             /* NCellsX == NCellsY */
  double dx, dy;
  const double* density;      /* C = nx*ny: stencil 1/rho and the dose */
  int32_t nm;
  const int32_t* degree;      /* nm */
  const double* Ax;           /* nm x nm  pn.Ax */
  const double* Az;           /* nm x nm  pn.Az */
  const double* absAx;        /* nm x nm  AbsAx */
  const double* absAz;        /* nm x nm  AbsAz */
  int32_t nq;
  const double* omega;        /* nq x 3   qReduced (:489) */
  const double* M;            /* nm x nq  MReduced (:490) */
  int32_t ne;                 /* energy grid points; steps 1..ne-1 */
  const double* dE;           /* ne-1: eTrafo[n] - eTrafo[n-1] of step n-1 (:527) */
  const double* stopping_mid; /* ne-1: csd.SMid[n-1] of step n-1 (:575) */
  int32_t nsigma;
  const double* sigma;        /* ne x nsigma, row-major: sigma_l(eGrid[n]) (:524) */
  int32_t nb;
  const int32_t* bcells;      /* the four edges of the grid (:528-537) */
  const double* bc_base;      /* nb x nq: PsiBeam without its energy factor */
  const double* bc_scale;     /* ne-1: energy factor of PsiBeam(energy[n], n-1), step n-1 */
  const double* psi0;         /* C x nq initial flux, SetupIC (:484); NULL: zeros */
} titus_fcs_problem_v1;

/* SolverCSD(settings) for the full solver: validate, build the block-sparse
 * operator plan, copy operands to device `device`. */
int titus_fcs_create(titus_fcs** out, const titus_fcs_problem_v1* problem, int32_t device);
/* One iteration of `for n=2:nEnergies` (:522-582); `step` (= n-1) must be the next. */
int titus_fcs_step(titus_fcs* ctx, int32_t step);
/* Remaining steps up to and including last_step (<= 0: to the end of the grid). */
int titus_fcs_solve(titus_fcs* ctx, int32_t last_step);
/* State (NULL pointers skipped): u C x nm (unscaled; the reference returns
 * 0.5*sqrt(gamma[1])*u), dose C, psi C x nq. */
int titus_fcs_get_state(titus_fcs* ctx, double* u, double* dose, double* psi);
/* Kernel timing with HIP events (on != 0) and per-kernel totals, as
 * titus_set_timing / titus_kernel_stats; names "bc", "sweep", "collide",
 * "dose". */
int titus_fcs_set_timing(titus_fcs* ctx, int32_t on);
int titus_fcs_kernel_stats(const titus_fcs* ctx, const char* name, double* total_ms,
                           int64_t* launches, double* bytes, double* flops);
/* Operator plan of the collided kernel (no reference counterpart): values[0]
 * column blocks per wavefront, [1] block groups, [2] nonzero 4 x 16 operator
 * blocks (matrix-core steps per 16 cells), [3] dense block count, [4] 1 when
 * chosen by timing at creation, [5] 16-cell tiles per wavefront, [6] 4-moment
 * K-steps per block.  Writes at most n values. */
int titus_fcs_get_plan(const titus_fcs* ctx, int32_t* values, int32_t n);
void titus_fcs_destroy(titus_fcs* ctx);

/* ---- device-pointer primitives --------------------------------------------
 * A structured grid whose density lives in device memory: cells
 * c = (i*ny + j)*nz + k as above (nz <= 1: a 2D grid, c = i*ny + j). */
typedef struct titus_grid_v1 {
  int32_t nx, ny, nz;
  double dx, dy, dz;
  const double* inv_density;  /* device, C = nx*ny*nz: 1 / rho */
} titus_grid_v1;

/* Galerkin products of a spatial basis in device memory, the
 * XL1xX = X'*obj.stencil.L1x*X, ... of SolverCSD.jl:1035-1038 (and of the
 * CUDA.jl variant, :1583-1584): out (device) receives 2 nd blocks of r x r,
 * column-major, in the order L1x, L1y, (L1z,) L2x, L2y, (L2z), each
 * X^T L_o X with the stencils of stencils.jl:12-130 acting on X / rho.
 * X: device, C x r with leading dimension ldx >= C, 1 <= r <= 96.  Runs on
 * `stream` (a hipStream_t; NULL: the default stream).  The reduction's
 * partial sums are allocated for the call and freed before it returns, so
 * the call waits for its stream. */
int titus_galerkin(const titus_grid_v1* grid, const double* X, int32_t ldx, int32_t r,
                   double* out, void* stream);

/* titus_galerkin without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as titus_galerkin, but the partial sums live
 * in a library-owned device array of 12 MiB per GPU (1024 x 6 x 256 doubles,
 * the bound of one reduction; every accepted size fits), kept until the
 * process ends, and the call does not wait for its stream.  All calls on one
 * GPU share that array: calls that overlap on one GPU (two streams, or two
 * threads) race and must be serialised by the caller. */
int titus_galerkin_noalloc(const titus_grid_v1* grid, const double* X, int32_t ldx, int32_t r,
                           double* out, void* stream);

const char* titus_last_error(void);
int32_t titus_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif /* TITUS_HIP_H */
