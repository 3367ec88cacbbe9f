// Small dense linear algebra on the host (FP64, column-major).
//
// The DLRA step needs factorizations of matrices whose sizes are set by the
// rank (r <= a few dozen) and the moment count (Nm <= a few hundred): the
// r x r / 2r x 2r SVDs of truncate! (code/SolverCSD.jl:3371), the SVD of the
// Nm x r angular factors (svd!/qr at SolverCSD.jl:1042,1073,1093,2180,2268),
// the Cholesky factors of the shifted CholeskyQR used for the tall spatial
// factors, and the canonical completion of rank-deficient bases.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace titus {

// A column-major matrix owned on the host.
struct HMat {
  int m = 0, n = 0;
  std::vector<double> a;
  HMat() = default;
  HMat(int m_, int n_) : m(m_), n(n_), a((std::size_t)m_ * n_, 0.0) {}
  double& operator()(int i, int j) { return a[(std::size_t)j * m + i]; }
  double operator()(int i, int j) const { return a[(std::size_t)j * m + i]; }
  double* col(int j) { return a.data() + (std::size_t)j * m; }
  const double* col(int j) const { return a.data() + (std::size_t)j * m; }
  static HMat eye(int m, int n);
};

// Compressed-row sparse matrix (the P_N streaming matrices have 2..64
// nonzeros per row out of Nm).
struct SpMat {
  int m = 0, n = 0;
  std::vector<int> rowptr, col;
  std::vector<double> val;
  static SpMat from_dense(const HMat& A);
};
// A * B for sparse A and dense B.
HMat spmm(const SpMat& A, const HMat& B);

// C = op(A) * op(B) with op = transpose when the flag is set.
HMat matmul(const HMat& A, const HMat& B, bool ta = false, bool tb = false);
HMat transpose(const HMat& A);
HMat hcat(const HMat& A, const HMat& B);
HMat cols(const HMat& A, int j0, int j1);   // columns [j0, j1)
double frob(const HMat& A);

// Thin SVD A = U diag(s) V^T of an m x n matrix (m >= n or m < n): U is m x k,
// V is n x k, k = min(m, n), singular values descending.  Householder QR
// followed by one-sided (Hestenes) Jacobi on the triangular factor: singular
// values carry errors of order eps * s_1, singular vectors of order
// eps * s_1 / gap, as LAPACK's gesvd.  The QR is column-pivoted and Jacobi
// runs on R^T (fewer sweeps on graded factors).  U and s are bit-identical
// with and without want_v; V is left empty without it (m >= n only).
void svd(const HMat& A, HMat& U, std::vector<double>& s, HMat& V, bool want_v = true);

// Thin SVD by Golub-Kahan-Reinsch: Householder bidiagonalization, then
// implicit-shift QR on the bidiagonal (the algorithm family of LAPACK's
// gesvd).  Same shapes and order as svd(); always forms V.  Its small singular
// directions carry the errors of a bidiagonalization (order eps * s_1), where
// the preconditioned Jacobi of svd() resolves them to high relative accuracy;
// truncate! follows the reference's LAPACK SVD, so it uses this one.
void svd_gk(const HMat& A, HMat& U, std::vector<double>& s, HMat& V);

// Cholesky A = R^T R of a symmetric positive definite matrix; returns false
// (and leaves R undefined) when a pivot is not positive.
bool cholesky_upper(const HMat& A, HMat& R);
// Inverse of an upper-triangular matrix.
HMat inv_upper(const HMat& R);

// Canonical completion.  Given an m x q block `base` with orthonormal columns
// and a list of candidate indices, choose `need` canonical vectors e_c,
// orthogonalised against `base` and against each other.
//   greedy == false: candidates in order, a candidate is taken when its
//                    residual norm^2 exceeds 1/2;
//   greedy == true : the candidate of largest residual norm^2 is taken each
//                    time; residuals within kCompletionTie (relative) of the
//                    largest are ties, won by the lowest index.
// Only the rows base[c, :] of the candidates are needed: the residual Gram
// matrix is I - Bc Bc^T.  Returns the chosen candidates and the upper factor
// Rinv (need x need) such that the new columns are
//     (E_chosen - base * Bc_chosen^T) * Rinv.
// Returns false when the candidates run out.
constexpr double kCompletionTie = 1e-6;
bool canonical_completion(const HMat& Bc, const std::vector<int>& cand_index,
                          int need, bool greedy, std::vector<int>& chosen,
                          HMat& Rinv);

// Orthonormal basis of range(A) minus span(against) (both m-row blocks),
// singular values <= tol * scale treated as zero (scale < 0: sigma_1 of the
// projected A), completed canonically from `candidates` (all rows when empty)
// to `ncols` columns.  Returns the basis (m x ncols) and its numerical rank.
// `rec` (without `against` only) receives what the adjoint of the basis
// update needs: the thin SVD A = U diag(sv) V^T, the kept rank p and the
// coefficients Tp (p x ncols) of the basis on the kept vectors U[:, :p].
struct OrthRecord {
  int p = 0;
  HMat U, V;
  std::vector<double> sv;
  HMat Tp;
};
// `chosen_out` receives the completion candidates taken, in column order.
HMat range_basis(const HMat& A, int ncols, const HMat* against, double scale,
                 double tol, bool greedy, const std::vector<int>& candidates,
                 int* rank_out = nullptr, OrthRecord* rec = nullptr,
                 std::vector<int>* chosen_out = nullptr);

// Adjoint of X1 = range_basis(A): Abar from X1bar.  X1 moves only through
// its kept singular subspace U_p, dX1 = (I - X1 X1^T) dU_p Tp (mod span X1);
// dU_p has the part (I - U U^T) dA V_p Sigma_p^-1 outside range(A) and, for
// each dropped pair j > p, u_j (s_i u_j^T dA v_i + s_j u_i^T dA v_j) /
// (s_i^2 - s_j^2).
HMat orth_bar(const OrthRecord& rec, const HMat& X1, const HMat& X1bar);

// SVD of a square matrix whose singular values <= tol * s_1 are set to 0 and
// whose corresponding singular vectors are completed canonically (greedy).
void svd_completed(const HMat& S, double tol, HMat& U, std::vector<double>& d, HMat& V);

}  // namespace titus
