// Device kernels of the DLRA energy step (gfx950) and their host launchers.
//
// Layout: every spatial block is column-major C x m with leading dimension
// `ld` (>= C): column a of block A starts at A + a*ld, cell c = i*ny + j.
// Small matrices (r x r, k x n) are column-major with their row count as
// leading dimension.
//
// Precision: spatial bases are stored in FP64.  The uncollided flux psi is
// stored in FP64 (plain form) or FP32 (mixed form); the streaming stencil,
// the Galerkin products and the psi contractions can compute in FP32
// (`fp32`), the Cholesky-QR passes and the dose accumulation always in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <vector>

namespace titus {
namespace dev {

// OP_ALL computes all stencil products of the grid at once: 2D, four
// consecutive m x n blocks L1x, L1y, L2x, L2y; 3D, six blocks L1x, L1y, L1z,
// L2x, L2y, L2z (dissipative operators first, then central ones).
enum StencilOp { OP_NONE = -1, OP_ALL = 4 };

// element type of a device block
enum class DT { F64, F32 };

// Structured grid; cells c = (i*ny + j)*nz + k, axis strides x: ny*nz,
// y: nz, z: 1.  A 2D grid has nz = 1.
struct Grid {
  int nx, ny, C;
  double dx, dy;
  const double* rho;            // C
  const double* irho;           // C, 1 / rho
  int nz = 1;
  double dz = 1.0;
  int ndim() const { return nz > 1 ? 3 : 2; }
};

// Launch plan of the uncollided sweep: `rows` per wavefront strip (2, 4, 8
// or 16), `groups` of ordinates (1..kMaxSweepGroups) that run side by side,
// and whether psi streams through non-temporal loads and stores.
// The l = 0 moment u0[c] = sum_q psi[c, q] m0[q] is kept as G partial sums
// u0p[g*C + c] (g < G = sweep_groups(...)), written by the sweep and
// refreshed at the boundary cells by launch_bc; the dose adds them up in
// order g = 0..G-1.
constexpr int kMaxSweepGroups = 8;
struct SweepPlan {
  int rows = 8;
  int groups = 1;
  bool nontemporal = false;   // psi streamed with non-temporal loads and stores
  bool limited = false;       // second-order limited flux (2D; no non-temporal form)
};
SweepPlan sweep_default(const Grid& g, int nq, DT tpsi);
std::vector<SweepPlan> sweep_candidates(const Grid& g, int nq, DT tpsi);
int sweep_groups(const Grid& g, int nq, const SweepPlan& plan);

// psi[bcells[b] + q*C] = bc_base[b + q*nb] * scale     (SolverCSD.jl:977-988),
// and u0p of the boundary cells (partial 0 = their moment, the others 0)
void launch_bc(hipStream_t st, void* psi, DT tpsi, int C, int nq, const int* bcells, int nb,
               const double* bc_base, double scale, const double* m0, double* u0p, int G);

// Analytic uncollided flux of the CUDA.jl variant (SolverCSD.jl:1549-1557) at
// transformed energy tau with the running int sigma: psi[c, q] set where both
// spatial normpdf factors of the beam are >= 1e-6, kept elsewhere; u0p (one
// partial) receives the l = 0 moment of the result.  2D grids.
void launch_analytic_psi(hipStream_t st, const Grid& g, void* psi, DT tpsi, int nq,
                         const double* omega, const double* xmid, const double* ymid, double x0,
                         double y0, double om1, double om3, double tau, double int_sigma,
                         const double* m0, double* u0p);

// dose[c] += coef * (X[c,:] . v + u0[c]) / rho[c]   (SolverCSD.jl:990-997), FP64
void launch_dose(hipStream_t st, const Grid& g, double* dose, const double* X, int ldx, int r,
                 const double* v, const double* u0p, int G, double coef);

// Uncollided upwind update with the boundary restore (SolverCSD.jl:388-463,
// 1000-1005): pout = (pin - dE * flux(pin/rho)) / denom on interior cells,
// pin / denom on ghost cells, pin itself for ordinate 0 of `restore` cells.
// plan.limited: the second-order limited flux of solveFlux! (SolverCSD.jl:
// 278-386) on cells 2..n-3 instead, pin / denom on the other cells.
// Computes in the storage type of psi; writes the moment partials of pout.
void launch_sweep(hipStream_t st, const Grid& g, const void* pin, void* pout, DT tpsi, int nq,
                  const double* omega, double dE, double denom, const unsigned char* restore,
                  const double* m0, double* u0p, const SweepPlan& plan);

// Streaming K-step (SolverCSD.jl:1020-1027) with the stencils applied matrix
// free: K[c,:] = X[c,:] Z0 - dE * sum_{d in {c,e,w,n,s}} X[d,:]/rho[d] Zd on
// interior cells and X[c,:] Z0 on ghost cells.  Z holds the six r x n blocks
// [Z0 Zc Ze Zw Zn Zs] side by side (r x 6n, column-major).
void launch_stream_k(hipStream_t st, const Grid& g, const double* X, int ldx, int r,
                     const double* Z, int n, double dE, double* K, int ldk, bool fp32);

// Tall-skinny A^T op(B) (m x n) with deterministic two-stage reduction.
// op = OP_NONE: B itself; OP_ALL: the 2*ndim stencil operators applied to B
// (the Galerkin products X^T L X of SolverCSD.jl:1035-1038).  `fp32`:
// compute and accumulate in FP32.  `partial` must hold tn_partial_size() doubles.
// Products run on the matrix cores (16 x 16 x 4 FP64 or FP32 MFMA tiles);
// op is OP_NONE or OP_ALL.
size_t tn_partial_size();
void launch_tn(hipStream_t st, const Grid& g, const void* A, DT ta, int lda, int m, const void* B,
               DT tb, int ldb, int n, int op, bool fp32, double* partial, double* out);
// FP64 convenience form
inline void launch_tn(hipStream_t st, const Grid& g, const double* A, int lda, int m,
                      const double* B, int ldb, int n, int op, double* partial, double* out) {
  launch_tn(st, g, A, DT::F64, lda, m, B, DT::F64, ldb, n, op, false, partial, out);
}

// Cholesky of a k x k SPD block on the device (k <= 96): G + s I = R^T R with
// the shift schedule of the host code (s = 0 first when allow_plain, then
// 1e-14 tr(G) x 10^i); writes R, R^-1 and status[0] (attempts, or -1).
void launch_chol_inv(hipStream_t st, const double* G, int k, bool allow_plain, double* R,
                     double* Rinv, double* status);

// out[c, b] = beta * out[c, b] + alpha * sum_a A[c, a] T[a, b] (out FP64);
// rows with zero_mask[c] != 0 are set to 0 afterwards (when zero_mask is not
// null).  `fp32`: products and sums in FP32.
void launch_gemm(hipStream_t st, int C, const void* A, DT ta, int lda, int k, const double* T,
                 int n, double alpha, double beta, double* out, int ldo,
                 const unsigned char* zero_mask, bool fp32);
inline void launch_gemm(hipStream_t st, int C, const double* A, int lda, int k, const double* T,
                        int n, double alpha, double beta, double* out, int ldo,
                        const unsigned char* zero_mask) {
  launch_gemm(st, C, A, DT::F64, lda, k, T, n, alpha, beta, out, ldo, zero_mask, false);
}

// The same two products with their coefficients (T: k x n, Z: the 2 nd + 2
// blocks of r x n) taken from host memory and carried in the kernel
// arguments, when they hold at most kArgMat values: no staging copy on the
// stream.  The host buffer may be reused as soon as the call returns.
constexpr int kArgMat = 448;
void launch_gemm_args(hipStream_t st, int C, const void* A, DT ta, int lda, int k,
                      const double* T_host, int n, double alpha, double beta, double* out, int ldo,
                      const unsigned char* zero_mask, bool fp32);
void launch_stream_k_args(hipStream_t st, const Grid& g, const double* X, int ldx, int r,
                          const double* Z_host, int n, double dE, double* K, int ldk, bool fp32);

// out[t + k*nrows] = A[rows[t] + k*lda], k < ncols.
void launch_gather_rows(hipStream_t st, const double* A, int lda, int ncols, const int* rows,
                        int nrows, double* out);

// A[rows[t] + cols[t]*lda] += vals[t] with packed = [rows, cols, vals] (n
// each, the indices stored as doubles).
void launch_scatter_add(hipStream_t st, double* A, int lda, const double* packed, int n);

// ---- adjoint (FP64; y = 1/rho) -------------------------------------------------
// K = X Z0 - dE sum_d [d interior] X[d,:] Zd: the transposed stencils, blocks
// laid out as for launch_stream_k, no 1/rho.
void launch_stream_kt(hipStream_t st, const Grid& g, const double* X, int ldx, int r,
                      const double* Z, int n, double dE, double* K, int ldk);
// out += alpha y o T (out may be null); ybar += alpha rowsum(V o T) (V may be null)
void launch_adj_combine(hipStream_t st, int C, int n, double* out, int ldo, const double* T, int ldt,
                        const double* V, int ldv, const double* rho, double alpha, double* ybar);
void launch_axpy(hipStream_t st, size_t n, double alpha, const double* x, double* y);
void launch_zero_rows(hipStream_t st, double* A, int lda, int ncols, const int* rows, int nrows);
// transpose of dose += a y (X v + psi m0) for the seed Db; writes g = a y Db
void launch_dose_adj(hipStream_t st, const Grid& g, const double* Db, double a, const double* X,
                     int ldx, int r, const double* v, const double* psi, int nq, const double* m0,
                     double* Xb, double* psib, double* ybar, double* gout);
// transpose of the 2D uncollided update (restore of ordinate 0 included):
// out = adjoint of psi before the sweep, ybar accumulates the density part
void launch_sweep_adj(hipStream_t st, const Grid& g, const double* G, const double* psi_a, int nq,
                      const double* omega, double dE, double denom, const unsigned char* restore,
                      double* out, double* ybar);

// ---- full-rank P_N solve (SolveFirstCollisionSource, SolverCSD.jl:465-586) -----
// Block-sparse moment operators of the collided streaming sum.  The four
// operators |Ax|, |Az|, Ax, Az (o = 0..3, paired with the stencils L1x, L1y,
// L2x, L2y) are cut into (4 ks) x 16 blocks: rows k = 4 ks kb .. 4 ks kb +
// 4 ks - 1, columns m = 16 mb .. 16 mb + 15.  Column blocks are dealt to
// `ngroups` groups of `nbw` (mb = g nbw + t).  mask[(g nkb + kb) 4 + o] bit t
// marks block (o, kb, g nbw + t) nonzero; the nonzero blocks of (g, kb) are
// stored from block index start[g (nkb + 1) + kb] on, ordered by o, then t,
// 64 ks values each (value s 64 + l = operator[4 ks kb + 4 s + l / 16,
// 16 mb + l % 16]), followed by 4 nbw zero blocks.
struct PnBlocks {
  int nm = 0, nkb = 0, nnb = 0, nbw = 0, ngroups = 0;
  int ct = 1;                 // 16-cell tiles per wavefront (1, or 2 when nbw <= 8)
  int ks = 1;                 // K-steps of 4 moments per block (1, or 2 when nbw <= 8)
  const unsigned* mask = nullptr;
  const int* start = nullptr;
  const double* pk = nullptr;
};
constexpr int kPnMaxNbw = 16;

// un[c, m] = (u[c, m] - dE (sum_o (L_o u)[c, :] A_o[:, m] - Dvec[m] (psi[c, :] M^T)[m]))
//            / (1 + dE (sigma_0 - Dvec[m]))                     on interior cells,
// 0 on ghost cells (SolverCSD.jl:556-571), Dvec[m] = sig[degree[m]]; the
// stencils act on u/rho (g.irho) as the reference's stencil matrices do
// (stencils.jl:26-128).  u, un: C x nm; psi: C x nq; M: nm x nq.
void launch_pn_collide(hipStream_t st, const Grid& g, const PnBlocks& b, const double* u, double* un,
                       const double* psi, int nq, const double* M, const int* degree,
                       const double* sig, double dE);

}  // namespace dev
}  // namespace titus
