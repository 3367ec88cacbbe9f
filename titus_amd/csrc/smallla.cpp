// Small dense linear algebra on the host; see smallla.h.
#include "smallla.h"

#include <array>
#include <cstring>
#include <utility>

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace titus {

static inline double dot(const double* a, const double* b, int n) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    s0 += a[i] * b[i];
    s1 += a[i + 1] * b[i + 1];
    s2 += a[i + 2] * b[i + 2];
    s3 += a[i + 3] * b[i + 3];
  }
  for (; i < n; ++i) s0 += a[i] * b[i];
  return (s0 + s1) + (s2 + s3);
}

HMat HMat::eye(int m, int n) {
  HMat I(m, n);
  for (int i = 0; i < std::min(m, n); ++i) I(i, i) = 1.0;
  return I;
}

SpMat SpMat::from_dense(const HMat& A) {
  SpMat S;
  S.m = A.m;
  S.n = A.n;
  S.rowptr.assign(A.m + 1, 0);
  for (int i = 0; i < A.m; ++i) {
    for (int j = 0; j < A.n; ++j)
      if (A(i, j) != 0.0) {
        S.col.push_back(j);
        S.val.push_back(A(i, j));
      }
    S.rowptr[i + 1] = (int)S.col.size();
  }
  return S;
}

namespace {
// Rows of A times B held row-major (Bt, width N): out row i accumulates whole
// rows of B over the row's nonzeros in order.  The width is a compile-time
// constant so the accumulator stays in registers and the row update is one
// unrolled vector loop.
template <int N>
void spmm_rows(const SpMat& A, const double* Bt, double* Ct) {
  for (int i = 0; i < A.m; ++i) {
    double acc[N] = {};
    for (int t = A.rowptr[i]; t < A.rowptr[i + 1]; ++t) {
      const double v = A.val[t];
      const double* b = Bt + (std::size_t)A.col[t] * N;
#pragma GCC unroll 64
      for (int j = 0; j < N; ++j) acc[j] += v * b[j];
    }
    std::memcpy(Ct + (std::size_t)i * N, acc, sizeof(acc));
  }
}

void spmm_rows_any(const SpMat& A, const double* Bt, double* Ct, int n) {
  std::vector<double> acc(n);
  for (int i = 0; i < A.m; ++i) {
    std::fill(acc.begin(), acc.end(), 0.0);
    for (int t = A.rowptr[i]; t < A.rowptr[i + 1]; ++t) {
      const double v = A.val[t];
      const double* b = Bt + (std::size_t)A.col[t] * n;
      for (int j = 0; j < n; ++j) acc[j] += v * b[j];
    }
    std::memcpy(Ct + (std::size_t)i * n, acc.data(), sizeof(double) * n);
  }
}

using SpmmRows = void (*)(const SpMat&, const double*, double*);
template <int... N>
constexpr std::array<SpmmRows, sizeof...(N) + 1> spmm_table(std::integer_sequence<int, N...>) {
  return {nullptr, &spmm_rows<N + 1>...};
}
constexpr int kSpmmMaxWidth = 48;
constexpr auto kSpmmRows = spmm_table(std::make_integer_sequence<int, kSpmmMaxWidth>{});
}  // namespace

HMat spmm(const SpMat& A, const HMat& B) {
  if (A.n != B.m) throw std::runtime_error("spmm: inner dimensions differ");
  // one pass over the nonzeros: each row of A combines whole rows of B (B held
  // row-major, so the inner loop runs over contiguous columns); the sum per
  // output entry runs over the row's nonzeros in order
  const int n = B.n;
  HMat C(A.m, n);
  std::vector<double> Bt((std::size_t)B.m * n), Ct((std::size_t)A.m * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < B.m; ++i) Bt[(std::size_t)i * n + j] = B(i, j);
  if (n >= 1 && n <= kSpmmMaxWidth) kSpmmRows[n](A, Bt.data(), Ct.data());
  else spmm_rows_any(A, Bt.data(), Ct.data(), n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < A.m; ++i) C(i, j) = Ct[(std::size_t)i * n + j];
  return C;
}

HMat matmul(const HMat& A, const HMat& B, bool ta, bool tb) {
  const int m = ta ? A.n : A.m, k = ta ? A.m : A.n;
  const int k2 = tb ? B.n : B.m, n = tb ? B.m : B.n;
  if (k != k2) throw std::runtime_error("matmul: inner dimensions differ");
  if (ta && !tb) {
    // A^T B: every entry is a dot product of two contiguous columns
    HMat C(m, n);
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) C(i, j) = dot(A.col(i), B.col(j), k);
    return C;
  }
  // transposed operands are materialised once, so every product runs as
  // column updates c += a_l * b_lj, vectorised over the rows; each entry is
  // still summed over l in ascending order
  HMat At, Bt;
  const HMat& Ao = ta ? (At = transpose(A)) : A;
  const HMat& Bo = tb ? (Bt = transpose(B)) : B;
  HMat C(m, n);
  for (int j = 0; j < n; ++j) {
    double* c = C.col(j);
    for (int l = 0; l < k; ++l) {
      const double b = Bo(l, j);
      if (b == 0.0) continue;
      const double* a = Ao.col(l);
      for (int i = 0; i < m; ++i) c[i] += a[i] * b;
    }
  }
  return C;
}

HMat transpose(const HMat& A) {
  HMat T(A.n, A.m);
  for (int j = 0; j < A.n; ++j)
    for (int i = 0; i < A.m; ++i) T(j, i) = A(i, j);
  return T;
}

HMat hcat(const HMat& A, const HMat& B) {
  if (A.n == 0) return B;
  if (B.n == 0) return A;
  if (A.m != B.m) throw std::runtime_error("hcat: row counts differ");
  HMat C(A.m, A.n + B.n);
  std::copy(A.a.begin(), A.a.end(), C.a.begin());
  std::copy(B.a.begin(), B.a.end(), C.a.begin() + A.a.size());
  return C;
}

HMat cols(const HMat& A, int j0, int j1) {
  HMat C(A.m, j1 - j0);
  std::copy(A.a.begin() + (size_t)j0 * A.m, A.a.begin() + (size_t)j1 * A.m, C.a.begin());
  return C;
}

double frob(const HMat& A) {
  double s = 0.0;
  for (double v : A.a) s += v * v;
  return std::sqrt(s);
}

// ---------------------------------------------------------------------------
// Q * B for the reflectors stored in QR (m x n), B has m rows (in place).
static void apply_q(const HMat& QR, const std::vector<double>& tau, HMat& B) {
  const int m = QR.m, n = QR.n;
  for (int j = n - 1; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    const double* v = QR.col(j);
    for (int k = 0; k < B.n; ++k) {
      double* b = B.col(k);
      double w = b[j] + dot(v + j + 1, b + j + 1, m - j - 1);
      w *= tau[j];
      b[j] -= w;
      for (int i = j + 1; i < m; ++i) b[i] -= w * v[i];
    }
  }
}

// One-sided Jacobi on a square n x n matrix B (overwritten by U diag(s));
// the rotations are accumulated in V only when want_v.
static void jacobi_square(HMat& B, HMat& V, bool want_v) {
  const int n = B.n;
  if (want_v) V = HMat::eye(n, n);
  const double tol = 1e-15;
  std::vector<double> nrm(n);
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int j = 0; j < n; ++j) nrm[j] = dot(B.col(j), B.col(j), B.m);
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double* bp = B.col(p);
        double* bq = B.col(q);
        const double alpha = nrm[p], beta = nrm[q];
        if (alpha == 0.0 || beta == 0.0) continue;
        const double gamma = dot(bp, bq, B.m);
        if (std::fabs(gamma) <= tol * std::sqrt(alpha * beta)) continue;
        rotated = true;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = c * t;
        for (int i = 0; i < B.m; ++i) {
          const double x = bp[i], y = bq[i];
          bp[i] = c * x - s * y;
          bq[i] = s * x + c * y;
        }
        if (want_v) {
          double* vp = V.col(p);
          double* vq = V.col(q);
          for (int i = 0; i < n; ++i) {
            const double x = vp[i], y = vq[i];
            vp[i] = c * x - s * y;
            vq[i] = s * x + c * y;
          }
        }
        // squared norms after the rotation (exact in exact arithmetic);
        // recomputed from the columns at the start of every sweep
        nrm[p] = alpha - t * gamma;
        nrm[q] = beta + t * gamma;
      }
    }
    if (!rotated) return;
  }
}

// Householder QR with column pivoting (the remaining column of largest norm
// first, ties to the lowest index): A P = Q R, perm[j] = original column of
// column j.  On return A holds R in its upper triangle and the reflector
// vectors (unit leading entry implied) below.
static void householder_qrcp(HMat& A, std::vector<double>& tau, std::vector<int>& perm) {
  const int m = A.m, n = A.n;
  tau.assign(n, 0.0);
  perm.resize(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::vector<double> nr(n);
  for (int j = 0; j < n; ++j) nr[j] = dot(A.col(j), A.col(j), m);
  for (int j = 0; j < n; ++j) {
    int p = j;
    for (int k = j + 1; k < n; ++k)
      if (nr[k] > nr[p]) p = k;
    if (p != j) {
      std::swap(perm[j], perm[p]);
      std::swap(nr[j], nr[p]);
      std::swap_ranges(A.col(j), A.col(j) + m, A.col(p));
    }
    double* a = A.col(j);
    const double xnorm2 = dot(a + j + 1, a + j + 1, m - j - 1);
    const double alpha = a[j];
    if (xnorm2 != 0.0) {
      const double nrm = std::sqrt(alpha * alpha + xnorm2);
      const double beta = alpha >= 0.0 ? -nrm : nrm;
      tau[j] = (beta - alpha) / beta;
      const double scal = 1.0 / (alpha - beta);
      for (int i = j + 1; i < m; ++i) a[i] *= scal;
      a[j] = beta;
      for (int k = j + 1; k < n; ++k) {
        double* b = A.col(k);
        double w = b[j] + dot(a + j + 1, b + j + 1, m - j - 1);
        w *= tau[j];
        b[j] -= w;
        for (int i = j + 1; i < m; ++i) b[i] -= w * a[i];
      }
    }
    // remaining norms recomputed (no downdating: the columns are short)
    for (int k = j + 1; k < n; ++k) nr[k] = dot(A.col(k) + j + 1, A.col(k) + j + 1, m - j - 1);
  }
}

// Preconditioned one-sided Jacobi (Drmac and Veselic, SIAM J. Matrix Anal.
// Appl. 29, 2008): A P = Q R with column pivoting, then one-sided Jacobi on
// R^T, X = R^T J with orthogonal columns.  R = J Sigma Ut^T gives A = (Q J)
// Sigma (P Ut)^T.  On the graded factors of a low-rank solve (rapidly
// decaying singular values) it converges in 3-5 sweeps instead of 6-8.
//
// U and s do not depend on want_v (the same operations run either way): the
// adjoint replays a solve segment from a checkpoint with want_v set, and its
// bases must be the forward's to the last bit, since the adjoint state
// handed across a checkpoint is expressed in the forward's basis.  A right
// vector of an exactly zero singular value is undetermined; those columns of
// V are an orthonormal completion of the others (their weight in A is zero).
void svd(const HMat& A0, HMat& U, std::vector<double>& s, HMat& V, bool want_v) {
  if (A0.m < A0.n) {
    HMat Ut, Vt;
    svd(transpose(A0), Ut, s, Vt, true);
    U = Vt;
    V = Ut;
    return;
  }
  const int m = A0.m, n = A0.n;
  HMat QR = A0;
  std::vector<double> tau;
  std::vector<int> perm;
  householder_qrcp(QR, tau, perm);
  HMat X(n, n);                                   // R^T
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) X(j, i) = QR(i, j);
  HMat J;
  jacobi_square(X, J, true);
  std::vector<double> sv(n);
  for (int j = 0; j < n; ++j) sv[j] = std::sqrt(dot(X.col(j), X.col(j), n));
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sv[a] > sv[b]; });
  HMat Ul(m, n);   // padded to m rows for apply_q
  V = want_v ? HMat(n, n) : HMat();
  s.assign(n, 0.0);
  int nz = 0;      // columns with a nonzero singular value (they sort first)
  for (int k = 0; k < n; ++k) {
    const int j = ord[k];
    s[k] = sv[j];
    for (int i = 0; i < n; ++i) Ul(i, k) = J(i, j);
    if (sv[j] > 0.0) {
      ++nz;
      if (want_v)
        for (int i = 0; i < n; ++i) V(perm[i], k) = X(i, j) / sv[j];
    }
  }
  apply_q(QR, tau, Ul);
  U = Ul;
  if (want_v && nz < n) {
    // completion: per column, the unit vector e_t with the largest part
    // outside the span so far (the first on ties), projected twice, normalized
    for (int k = nz; k < n; ++k) {
      int best = 0;
      double bn = -1.0;
      for (int t = 0; t < n; ++t) {
        double r = 1.0;
        for (int c = 0; c < k; ++c) r -= V(t, c) * V(t, c);
        if (r > bn) { bn = r; best = t; }
      }
      double* v = V.col(k);
      std::fill(v, v + n, 0.0);
      v[best] = 1.0;
      for (int pass = 0; pass < 2; ++pass)
        for (int c = 0; c < k; ++c) {
          const double w = dot(V.col(c), v, n);
          for (int i = 0; i < n; ++i) v[i] -= w * V(i, c);
        }
      const double nv = std::sqrt(dot(v, v, n));
      for (int i = 0; i < n; ++i) v[i] /= nv;
    }
  }
}

bool cholesky_upper(const HMat& A, HMat& R) {
  const int n = A.n;
  R = HMat(n, n);
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i <= j; ++i) {
      double s = A(i, j);
      for (int k = 0; k < i; ++k) s -= R(k, i) * R(k, j);
      if (i == j) {
        if (!(s > 0.0)) return false;
        R(j, j) = std::sqrt(s);
      } else {
        R(i, j) = s / R(i, i);
      }
    }
  }
  return true;
}

HMat inv_upper(const HMat& R) {
  const int n = R.n;
  HMat X(n, n);
  for (int j = 0; j < n; ++j) {
    X(j, j) = 1.0 / R(j, j);
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += R(i, k) * X(k, j);
      X(i, j) = -s / R(i, i);
    }
  }
  return X;
}

bool canonical_completion(const HMat& Bc, const std::vector<int>& cand_index, int need,
                          bool greedy, std::vector<int>& chosen, HMat& Rinv) {
  // Bc: ncand x q rows of the orthonormal block at the candidates.
  // H(a, b) = [c_a == c_b] - Bc(a,:) . Bc(b,:).  Incremental Cholesky over the
  // chosen candidates; L[s][t] = v_s[c_t].
  const int nc = Bc.m, q = Bc.n;
  chosen.clear();
  std::vector<int> pos;            // positions (rows of Bc) of chosen candidates
  std::vector<std::vector<double>> L;   // L[s][t] for all candidate rows t
  std::vector<double> piv(nc);
  auto H = [&](int a, int b) {
    double s = (cand_index[a] == cand_index[b]) ? 1.0 : 0.0;
    for (int k = 0; k < q; ++k) s -= Bc(a, k) * Bc(b, k);
    return s;
  };
  auto pivot = [&](int t) {
    double p = H(t, t);
    for (size_t s = 0; s < L.size(); ++s) p -= L[s][t] * L[s][t];
    return p;
  };
  auto accept = [&](int t, double p) {
    const double nrm = std::sqrt(p);
    std::vector<double> row(nc, 0.0);
    for (int u = 0; u < nc; ++u) {
      double h = H(t, u);
      for (size_t s = 0; s < L.size(); ++s) h -= L[s][t] * L[s][u];
      row[u] = h / nrm;
    }
    L.push_back(std::move(row));
    pos.push_back(t);
    chosen.push_back(cand_index[t]);
  };
  std::vector<char> used(nc, 0);
  if (!greedy) {
    for (int t = 0; t < nc && (int)chosen.size() < need; ++t) {
      const double p = pivot(t);
      if (p > 0.5) { used[t] = 1; accept(t, p); }
    }
  }
  // greedy selection; also the fallback of the in-order rule when the
  // candidates run out.  Residuals within kCompletionTie (relative) of the
  // largest count as tied and the lowest index wins, so that perturbations of
  // the basis far below the tie width (FP32 arithmetic upstream) cannot change
  // which canonical vector completes it.
  {
    std::vector<double> pv(nc);
    while ((int)chosen.size() < need) {
      double pmax = 1e-8;
      for (int t = 0; t < nc; ++t) {
        pv[t] = used[t] ? -1.0 : pivot(t);
        if (pv[t] > pmax) pmax = pv[t];
      }
      int best = -1;
      for (int t = 0; t < nc; ++t)
        if (!used[t] && pv[t] > 1e-8 && pv[t] >= pmax * (1.0 - kCompletionTie)) { best = t; break; }
      if (best < 0) return false;
      used[best] = 1;
      accept(best, pv[best]);
    }
  }
  const int k = (int)chosen.size();
  HMat R(k, k);
  for (int s = 0; s < k; ++s)
    for (int s2 = s; s2 < k; ++s2) R(s, s2) = L[s][pos[s2]];
  Rinv = inv_upper(R);
  return true;
}

HMat range_basis(const HMat& A, int ncols, const HMat* against, double scale, double tol,
                 bool greedy, const std::vector<int>& candidates, int* rank_out, OrthRecord* rec,
                 std::vector<int>* chosen_out) {
  const int m = A.m;
  if (rec && against && against->n > 0)
    throw std::runtime_error("range_basis: no adjoint record with a basis to project out");
  HMat Ap = A;
  if (against && against->n > 0) {
    for (int pass = 0; pass < 2; ++pass) {
      HMat P = matmul(*against, Ap, true, false);
      HMat G = matmul(*against, P);
      for (size_t i = 0; i < Ap.a.size(); ++i) Ap.a[i] -= G.a[i];
    }
  }
  HMat U, V;
  std::vector<double> sv;
  int p = 0;
  if (Ap.n > 0) {
    svd(Ap, U, sv, V, rec != nullptr);
    if (scale < 0.0) scale = sv.empty() ? 0.0 : sv[0];
    if (scale > 0.0)
      for (double x : sv) if (x > tol * scale) ++p;
  }
  p = std::min(p, ncols);
  HMat B = p > 0 ? cols(U, 0, p) : HMat(m, 0);
  if (rank_out) *rank_out = p;
  if (chosen_out) chosen_out->clear();
  const int need = ncols - p;
  if (rec) {
    rec->p = p;
    rec->U = U;
    rec->V = V;
    rec->sv = sv;
    rec->Tp = HMat(p, ncols);
    for (int i = 0; i < p; ++i) rec->Tp(i, i) = 1.0;
  }
  if (need == 0) return B;
  HMat base = (against && against->n > 0) ? hcat(*against, B) : B;
  std::vector<int> cand = candidates;
  if (cand.empty()) {
    cand.resize(m);
    std::iota(cand.begin(), cand.end(), 0);
  }
  HMat Bc((int)cand.size(), base.n);
  for (size_t t = 0; t < cand.size(); ++t)
    for (int k = 0; k < base.n; ++k) Bc((int)t, k) = base(cand[t], k);
  std::vector<int> chosen;
  HMat Rinv;
  if (!canonical_completion(Bc, cand, need, greedy, chosen, Rinv))
    throw std::runtime_error("canonical completion ran out of candidates");
  if (chosen_out) *chosen_out = chosen;
  if (rec && p > 0) {
    // completion (E - U_p U_p[chosen]^T) Rinv: coefficients -U_p[chosen]^T Rinv
    HMat Uch(need, p);
    for (int s = 0; s < need; ++s)
      for (int k = 0; k < p; ++k) Uch(s, k) = B(chosen[s], k);
    HMat T = matmul(Uch, Rinv, true, false);
    for (int s = 0; s < need; ++s)
      for (int k = 0; k < p; ++k) rec->Tp(k, p + s) = -T(k, s);
  }
  // Y = E_chosen - base * Bc_chosen^T, completion = Y * Rinv
  HMat Y(m, need);
  for (int s = 0; s < need; ++s) {
    const int c = chosen[s];
    for (int i = 0; i < m; ++i) {
      double v = 0.0;
      for (int k = 0; k < base.n; ++k) v += base(i, k) * base(c, k);
      Y(i, s) = (i == c ? 1.0 : 0.0) - v;
    }
  }
  return hcat(B, matmul(Y, Rinv));
}

HMat orth_bar(const OrthRecord& rec, const HMat& X1, const HMat& X1bar) {
  const int m = rec.U.m, k = (int)rec.sv.size(), p = rec.p;
  HMat Ab(m, rec.V.m);
  if (p == 0) return Ab;
  // adjoint of U_p, projected off span(X1)
  HMat Ub = matmul(X1bar, rec.Tp, false, true);                 // m x p
  HMat h = matmul(X1, Ub, true, false);
  HMat Xh = matmul(X1, h);
  for (size_t i = 0; i < Ub.a.size(); ++i) Ub.a[i] -= Xh.a[i];
  // (I - U U^T) Ub Sigma_p^-1 V_p^T
  HMat Y = Ub;
  for (int c = 0; c < p; ++c)
    for (int i = 0; i < m; ++i) Y(i, c) /= rec.sv[c];
  HMat UtY = matmul(rec.U, Y, true, false);                     // k x p
  HMat UUtY = matmul(rec.U, UtY);
  for (size_t i = 0; i < Y.a.size(); ++i) Y.a[i] -= UUtY.a[i];
  HMat Vp = cols(rec.V, 0, p);
  Ab = matmul(Y, Vp, false, true);
  if (k > p) {
    // beta = U_d^T Ub; coupling of each kept i with each dropped j
    HMat UtUb = matmul(rec.U, Ub, true, false);                 // k x p
    HMat B1(k - p, p), B2(p, k - p);
    for (int i = 0; i < p; ++i)
      for (int j = p; j < k; ++j) {
        const double si = rec.sv[i], sj = rec.sv[j];
        const double f = UtUb(j, i) / (si * si - sj * sj);
        B1(j - p, i) = f * si;
        B2(i, j - p) = f * sj;
      }
    HMat Ud = cols(rec.U, p, k), Up = cols(rec.U, 0, p), Vd = cols(rec.V, p, k);
    HMat t1 = matmul(matmul(Ud, B1), Vp, false, true);
    HMat t2 = matmul(matmul(Up, B2), Vd, false, true);
    for (size_t i = 0; i < Ab.a.size(); ++i) Ab.a[i] += t1.a[i] + t2.a[i];
  }
  return Ab;
}

// ---------------------------------------------------------------------------
// Golub-Kahan-Reinsch SVD (Golub and Reinsch, Numer. Math. 14, 1970; Golub
// and Van Loan, Matrix Computations, sec. 8.6): Householder bidiagonalization
// A = Q B P^T (column reflector, then row reflector, per column), then
// implicit Wilkinson-shift QR sweeps on the bidiagonal B with deflation, the
// rotations accumulated into U = Q and V = P.  The same algorithm family as
// LAPACK's gesvd/gesdd (reflectors, then a bidiagonal SVD).

namespace {

// Householder reflector H = I - tau v v^T with H [alpha; x] = [beta; 0],
// v = [1; x / (alpha - beta)]; x is overwritten by v's tail.  tau = 0 (H = I,
// beta = alpha) when x is zero.
double make_reflector(double& alpha, double* x, int nx, int stride) {
  double xn2 = 0.0;
  for (int i = 0; i < nx; ++i) xn2 += x[(size_t)i * stride] * x[(size_t)i * stride];
  if (xn2 == 0.0) return 0.0;
  const double nrm = std::sqrt(alpha * alpha + xn2);
  const double beta = alpha >= 0.0 ? -nrm : nrm;
  const double tau = (beta - alpha) / beta;
  const double scal = 1.0 / (alpha - beta);
  for (int i = 0; i < nx; ++i) x[(size_t)i * stride] *= scal;
  alpha = beta;
  return tau;
}

// B = diag(d) + superdiag(e) (e[i] couples rows/columns i and i + 1) with
// A = U B V^T; on return d >= 0 holds the singular values (unsorted) and U,
// V the accumulated rotations.
void bidiag_qr(std::vector<double>& d, std::vector<double>& e, HMat& U, HMat& V) {
  const int n = (int)d.size();
  const double eps = std::numeric_limits<double>::epsilon();
  double anorm = 0.0;
  for (int i = 0; i < n; ++i)
    anorm = std::max(anorm, std::fabs(d[i]) + (i + 1 < n ? std::fabs(e[i]) : 0.0));
  const double small = eps * anorm;
  // columns a, b of M <- (c a + s b, -s a + c b)
  auto rot = [](HMat& M, int a, int b, double c, double s) {
    double* x = M.col(a);
    double* y = M.col(b);
    for (int i = 0; i < M.m; ++i) {
      const double xa = x[i], yb = y[i];
      x[i] = c * xa + s * yb;
      y[i] = -s * xa + c * yb;
    }
  };
  for (int hi = n - 1; hi >= 0; --hi) {
    for (int iter = 0;; ++iter) {
      if (iter > 100) throw std::runtime_error("svd: bidiagonal QR did not converge");
      // the unreduced block lo..hi: split where a coupling is negligible
      // against its two diagonal neighbours (or the matrix); a negligible
      // diagonal entry above the block is rotated out first
      int lo = hi;
      bool zero_diag = false;
      while (lo > 0) {
        const double el = std::fabs(e[lo - 1]);
        if (el <= eps * (std::fabs(d[lo - 1]) + std::fabs(d[lo])) || el <= small) {
          e[lo - 1] = 0.0;
          break;
        }
        if (std::fabs(d[lo - 1]) <= small) {
          zero_diag = true;
          break;
        }
        --lo;
      }
      if (zero_diag) {
        // d[k] ~ 0 with e[k] != 0 (k = lo - 1): left rotations of row k
        // against rows lo..hi move e[k] along the row and out of the block
        const int k = lo - 1;
        double f = e[k];
        e[k] = 0.0;
        d[k] = 0.0;
        for (int i = lo; i <= hi; ++i) {
          const double g = d[i];
          const double h = std::hypot(f, g);
          const double c = g / h, s = f / h;
          d[i] = h;
          // rows (i, k): row_i <- c row_i + s row_k, row_k <- -s row_i + c row_k
          rot(U, i, k, c, s);
          if (i < hi) {
            f = -s * e[i];
            e[i] = c * e[i];
            if (std::fabs(f) <= small) break;
          }
        }
        continue;
      }
      if (lo == hi) {
        if (d[hi] < 0.0) {
          d[hi] = -d[hi];
          double* v = V.col(hi);
          for (int i = 0; i < V.m; ++i) v[i] = -v[i];
        }
        break;
      }
      // Wilkinson shift: the eigenvalue of the trailing 2 x 2 of B^T B
      // nearer its last diagonal entry
      const double t11 = d[hi - 1] * d[hi - 1] + (hi - 1 > lo ? e[hi - 2] * e[hi - 2] : 0.0);
      const double t12 = d[hi - 1] * e[hi - 1];
      const double t22 = d[hi] * d[hi] + e[hi - 1] * e[hi - 1];
      const double del = 0.5 * (t11 - t22);
      const double den = del + (del >= 0.0 ? 1.0 : -1.0) * std::hypot(del, t12);
      const double mu = den != 0.0 ? t22 - t12 * t12 / den : t22;
      // chase the bulge from lo to hi
      double y = d[lo] * d[lo] - mu;
      double z = d[lo] * e[lo];
      for (int k = lo; k < hi; ++k) {
        // right rotation of columns k, k + 1: [y z] -> [r 0]
        double r = std::hypot(y, z);
        double c = r != 0.0 ? y / r : 1.0, s = r != 0.0 ? z / r : 0.0;
        if (k > lo) e[k - 1] = r;
        const double dk = c * d[k] + s * e[k];
        const double ek = -s * d[k] + c * e[k];
        const double bulge_lo = s * d[k + 1];           // at (k + 1, k)
        const double dk1 = c * d[k + 1];
        rot(V, k, k + 1, c, s);
        // left rotation of rows k, k + 1: [dk; bulge_lo] -> [r; 0]
        r = std::hypot(dk, bulge_lo);
        c = r != 0.0 ? dk / r : 1.0;
        s = r != 0.0 ? bulge_lo / r : 0.0;
        d[k] = r;
        e[k] = c * ek + s * dk1;
        d[k + 1] = -s * ek + c * dk1;
        rot(U, k, k + 1, c, s);
        if (k + 1 < hi) {
          y = e[k];
          z = s * e[k + 1];                              // bulge at (k, k + 2)
          e[k + 1] = c * e[k + 1];
        }
      }
    }
  }
}

}  // namespace

void svd_gk(const HMat& A0, HMat& U, std::vector<double>& s, HMat& V) {
  if (A0.m < A0.n) {
    HMat Ut, Vt;
    svd_gk(transpose(A0), Ut, s, Vt);
    U = Vt;
    V = Ut;
    return;
  }
  const int m = A0.m, n = A0.n;
  HMat A = A0;
  std::vector<double> tq(n, 0.0), tp(n, 0.0), d(n, 0.0), e(std::max(0, n - 1), 0.0);
  for (int j = 0; j < n; ++j) {
    // column reflector: zero A(j+1:m, j)
    double* a = A.col(j);
    tq[j] = make_reflector(a[j], a + j + 1, m - j - 1, 1);
    d[j] = a[j];
    if (tq[j] != 0.0)
      for (int k = j + 1; k < n; ++k) {
        double* b = A.col(k);
        const double w = tq[j] * (b[j] + dot(a + j + 1, b + j + 1, m - j - 1));
        b[j] -= w;
        for (int i = j + 1; i < m; ++i) b[i] -= w * a[i];
      }
    if (j + 1 < n) {
      // row reflector: zero A(j, j+2:n)
      double& alpha = A(j, j + 1);
      tp[j] = make_reflector(alpha, &A(j, std::min(j + 2, n - 1)), n - j - 2, m);
      e[j] = alpha;
      if (tp[j] != 0.0)
        for (int i = j + 1; i < m; ++i) {
          double w = A(i, j + 1);
          for (int k = j + 2; k < n; ++k) w += A(i, k) * A(j, k);
          w *= tp[j];
          A(i, j + 1) -= w;
          for (int k = j + 2; k < n; ++k) A(i, k) -= w * A(j, k);
        }
    }
  }
  // U = H_0 ... H_{n-1} I(m x n), V = G_0 ... G_{n-2} (G_j acts on j+1..n-1)
  HMat Uq(m, n), Vp = HMat::eye(n, n);
  for (int j = 0; j < n; ++j) Uq(j, j) = 1.0;
  for (int j = n - 1; j >= 0; --j) {
    if (tq[j] == 0.0) continue;
    const double* v = A.col(j);
    for (int k = j; k < n; ++k) {
      double* b = Uq.col(k);
      const double w = tq[j] * (b[j] + dot(v + j + 1, b + j + 1, m - j - 1));
      b[j] -= w;
      for (int i = j + 1; i < m; ++i) b[i] -= w * v[i];
    }
  }
  for (int j = n - 3; j >= 0; --j) {
    if (tp[j] == 0.0) continue;
    // v = [1; A(j, j+2:n)] on rows j+1..n-1
    for (int k = j + 1; k < n; ++k) {
      double* b = Vp.col(k);
      double w = b[j + 1];
      for (int i = j + 2; i < n; ++i) w += A(j, i) * b[i];
      w *= tp[j];
      b[j + 1] -= w;
      for (int i = j + 2; i < n; ++i) b[i] -= w * A(j, i);
    }
  }
  bidiag_qr(d, e, Uq, Vp);
  // descending order (stable: equal values keep their positions)
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] > d[b]; });
  U = HMat(m, n);
  V = HMat(n, n);
  s.assign(n, 0.0);
  for (int k = 0; k < n; ++k) {
    s[k] = d[ord[k]];
    std::copy(Uq.col(ord[k]), Uq.col(ord[k]) + m, U.col(k));
    std::copy(Vp.col(ord[k]), Vp.col(ord[k]) + n, V.col(k));
  }
}

void svd_completed(const HMat& S, double tol, HMat& U, std::vector<double>& d, HMat& V) {
  const int n = S.n;
  HMat Us, Vs;
  std::vector<double> sv;
  svd_gk(S, Us, sv, Vs);
  int p = 0;
  if (!sv.empty() && sv[0] > 0.0)
    for (double x : sv) if (x > tol * sv[0]) ++p;
  HMat Up = cols(Us, 0, p), Vp = cols(Vs, 0, p);
  std::vector<int> none;
  HMat Ec(n, 0);
  if (p < n) {
    U = hcat(Up, range_basis(Ec, n - p, &Up, 1.0, tol, true, none));
    V = hcat(Vp, range_basis(Ec, n - p, &Vp, 1.0, tol, true, none));
  } else {
    U = Up;
    V = Vp;
  }
  d.assign(n, 0.0);
  for (int i = 0; i < p; ++i) d[i] = sv[i];
}

}  // namespace titus
