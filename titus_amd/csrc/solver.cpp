// C-ABI and host orchestration of the DLRA energy step (see include/titus_hip.h).
//
// Division of work: every operation whose size grows with the cell count C
// (boundary source, dose, uncollided sweep, streaming stencils, Galerkin
// products, tall-skinny Gram and projection products, basis updates) runs on
// the device; the r x r, 2r x 2r and Nm x r algebra (angular factor, S
// updates, SVDs, Cholesky factors, completion) runs on the host in FP64.
//
// Spatial bases.  svd!(K).U (code/SolverCSD.jl:1029,1085) and
// qr([X K]).Q (code/SolverCSD.jl:2173,2262) only matter through the subspace
// they span, except where K is rank deficient and LAPACK fills the missing
// columns with rounding-determined directions.  The basis here is the range
// of K -- singular values above rank_tol times the scale -- completed by
// canonical vectors of interior cells (which is what LAPACK returns for an
// exactly zero block).  The range comes from a shifted CholeskyQR3 of K on the
// device (three Gram/Cholesky passes; backward error O(eps |K|)) followed by
// the SVD of the k x k triangular factor on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/titus_hip.h"
#include "kernels.h"
#include "smallla.h"

using titus::HMat;
namespace dev = titus::dev;

namespace {

thread_local std::string g_last_error;

struct TitusError : std::runtime_error {
  int code;
  TitusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      throw TitusError(TITUS_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
T* dalloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  return static_cast<T*>(p);
}

// Kernel timing with HIP events on the context's stream.
struct Timing {
  bool enabled = false;
  struct Rec { int id; hipEvent_t a, b; double bytes, flops; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, int> ids;
  std::vector<std::string> names;
  std::vector<double> total_ms, total_bytes, total_flops;
  std::vector<int64_t> launches;

  hipEvent_t ev() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    return e;
  }
  int id(const std::string& n) {
    auto it = ids.find(n);
    if (it != ids.end()) return it->second;
    int k = (int)names.size();
    ids[n] = k; names.push_back(n); total_ms.push_back(0.0); launches.push_back(0);
    total_bytes.push_back(0.0); total_flops.push_back(0.0);
    return k;
  }
  void resolve() {
    for (auto& r : pending) {
      float ms = 0.f;
      HIPCHK(hipEventSynchronize(r.b));
      HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
      total_ms[r.id] += ms;
      launches[r.id] += 1;
      total_bytes[r.id] += r.bytes;
      total_flops[r.id] += r.flops;
      pool.push_back(r.a); pool.push_back(r.b);
    }
    pending.clear();
  }
  ~Timing() {
    for (auto& r : pending) { hipEventDestroy(r.a); hipEventDestroy(r.b); }
    for (auto e : pool) hipEventDestroy(e);
  }
};

}  // namespace

struct titus_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  // problem (host copy)
  int nx = 0, ny = 0, C = 0, nm = 0, nq = 0, ne = 0, nsig = 0, integrator = 0, precision = 0;
  double dx = 0, dy = 0, dE = 0, theta = 0, theta_index = 0, rank_tol = 1e-7;
  int rank_cap = 0, capx = 0, capw = 0;
  std::vector<double> stopping, sigma, bc_scale;
  std::vector<int> degree, cand;
  HMat Ax, Az, aAx, aAz, M;
  // device buffers
  double *rho = nullptr, *omega = nullptr, *bc_base = nullptr, *m0 = nullptr;
  double *psi = nullptr, *psi2 = nullptr, *dose = nullptr;
  double *X = nullptr, *Xn = nullptr, *K = nullptr, *Q1 = nullptr, *Q2 = nullptr;
  double *partial = nullptr, *arena = nullptr, *upl = nullptr, *svals = nullptr;
  int *bcells = nullptr, *cand_d = nullptr, *srows = nullptr, *scols = nullptr;
  unsigned char *bmask = nullptr, *restore = nullptr;
  size_t arena_n = 0, upl_n = 0, scat_n = 0;
  dev::Grid g{};
  int nb = 0;
  // state
  HMat S, W;
  int rank = 0, steps_done = 0;
  std::vector<int> ranks;
  std::vector<double> etas;
  Timing timing;

  ~titus_ctx() {
    void* ptrs[] = {rho, omega, bc_base, m0, psi, psi2, dose, X, Xn, K, Q1, Q2, partial, arena,
                    upl, svals, bcells, cand_d, srows, scols, bmask, restore};
    for (void* p : ptrs) if (p) hipFree(p);
    if (st) hipStreamDestroy(st);
  }

  // ---- timing -------------------------------------------------------------
  // `bytes` / `flops`: algorithmic HBM traffic and arithmetic of the launch
  // (compulsory operand reads and result writes; reduction partials excluded)
  template <class F>
  void timed(const char* name, double bytes, double flops, F&& f) {
    if (!timing.enabled) { f(); return; }
    const int id = timing.id(name);
    hipEvent_t a = timing.ev(), b = timing.ev();
    HIPCHK(hipEventRecord(a, st));
    f();
    HIPCHK(hipEventRecord(b, st));
    timing.pending.push_back({id, a, b, bytes, flops});
  }

  // ---- transfers ------------------------------------------------------------
  double* upload(const HMat& h, size_t offset = 0) {
    if (offset + h.a.size() > upl_n) throw TitusError(TITUS_E_ARG, "upload area too small");
    double* d = upl + offset;
    if (!h.a.empty())
      HIPCHK(hipMemcpyAsync(d, h.a.data(), h.a.size() * sizeof(double), hipMemcpyHostToDevice, st));
    return d;
  }
  double* upload_vec(const std::vector<double>& v, size_t offset) {
    if (offset + v.size() > upl_n) throw TitusError(TITUS_E_ARG, "upload area too small");
    double* d = upl + offset;
    if (!v.empty())
      HIPCHK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, st));
    return d;
  }
  // copy `n` doubles of the arena (from `off`) to the host and wait
  std::vector<double> fetch(size_t off, size_t n) {
    std::vector<double> h(n);
    if (n) HIPCHK(hipMemcpyAsync(h.data(), arena + off, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return h;
  }
  static HMat as_mat(const std::vector<double>& v, size_t off, int m, int n) {
    HMat A(m, n);
    std::copy(v.begin() + off, v.begin() + off + (size_t)m * n, A.a.begin());
    return A;
  }

  // ---- device building blocks --------------------------------------------
  // arena[off : off + m*n] = A^T op(B)
  void tn(const double* A, int m, const double* B, int n, int op, size_t off) {
    if (off + (size_t)m * n > arena_n) throw TitusError(TITUS_E_ARG, "arena too small");
    const double by = 8.0 * C * (A == B ? m : m + n) + (op == dev::OP_NONE ? 0.0 : 8.0 * C);
    timed(op == dev::OP_NONE ? "gram" : "galerkin", by, 2.0 * C * m * n, [&] {
      dev::launch_tn(st, g, A, C, m, B, C, n, op, partial, arena + off);
    });
  }
  // [Xa^T L1x Xb, Xa^T L1y Xb, Xa^T L2x Xb, Xa^T L2y Xb]
  std::vector<HMat> galerkin_x(const double* Xa, int ma, const double* Xb, int mb) {
    const size_t blk = (size_t)ma * mb;
    for (int op = 0; op < 4; ++op) tn(Xa, ma, Xb, mb, op, op * blk);
    std::vector<double> h = fetch(0, 4 * blk);
    std::vector<HMat> out;
    for (int op = 0; op < 4; ++op) out.push_back(as_mat(h, op * blk, ma, mb));
    return out;
  }
  HMat tn_fetch(const double* A, int m, const double* B, int n) {
    tn(A, m, B, n, dev::OP_NONE, 0);
    return as_mat(fetch(0, (size_t)m * n), 0, m, n);
  }
  void gemm(const double* A, int k, const HMat& T, double alpha, double beta, double* out,
            const unsigned char* mask, size_t upl_off = 0) {
    const double* Td = upload(T, upl_off);
    timed("gemm", 8.0 * C * (k + T.n * (beta != 0.0 ? 2 : 1)), 2.0 * C * k * T.n, [&] {
      dev::launch_gemm(st, C, A, C, k, Td, T.n, alpha, beta, out, C, mask);
    });
  }

  // ---- host building blocks (angular side) -----------------------------------
  std::vector<HMat> galerkin_w(const HMat& Wa, const HMat& Wb) const {
    return {titus::matmul(Wa, titus::matmul(aAx, Wb), true), titus::matmul(Wa, titus::matmul(aAz, Wb), true),
            titus::matmul(Wa, titus::matmul(Ax, Wb), true), titus::matmul(Wa, titus::matmul(Az, Wb), true)};
  }
  // L - dE (Ax L XL2xX^T + Az L XL2yX^T + |Ax| L XL1xX^T + |Az| L XL1yX^T)
  HMat stream_L(const HMat& L, const std::vector<HMat>& XLX) const {
    HMat out = L;
    const HMat* A[4] = {&aAx, &aAz, &Ax, &Az};
    for (int k : {2, 3, 0, 1}) {
      HMat t = titus::matmul(titus::matmul(*A[k], L), XLX[k], false, true);
      for (size_t i = 0; i < out.a.size(); ++i) out.a[i] -= dE * t.a[i];
    }
    return out;
  }
  HMat stream_S(const HMat& S0, const std::vector<HMat>& XLX, const std::vector<HMat>& WAW) const {
    HMat out = S0;
    for (int k : {2, 3, 0, 1}) {
      HMat t = titus::matmul(titus::matmul(XLX[k], S0), WAW[k]);
      for (size_t i = 0; i < out.a.size(); ++i) out.a[i] -= dE * t.a[i];
    }
    return out;
  }
  // Z = [S, S(M1x/dx + M1y/dy), S(M2x-M1x)/(2dx), S(-M2x-M1x)/(2dx),
  //      S(M2y-M1y)/(2dy), S(-M2y-M1y)/(2dy)] with Mk = W^T A_k W
  HMat stream_Z(const HMat& S0, const std::vector<HMat>& WAW) const {
    const int r = S0.m, n = S0.n;
    const HMat &M1x = WAW[0], &M1y = WAW[1], &M2x = WAW[2], &M2y = WAW[3];
    HMat comb[5] = {HMat(n, n), HMat(n, n), HMat(n, n), HMat(n, n), HMat(n, n)};
    for (size_t i = 0; i < (size_t)n * n; ++i) {
      comb[0].a[i] = M1x.a[i] / dx + M1y.a[i] / dy;
      comb[1].a[i] = (M2x.a[i] - M1x.a[i]) / (2.0 * dx);
      comb[2].a[i] = (-M2x.a[i] - M1x.a[i]) / (2.0 * dx);
      comb[3].a[i] = (M2y.a[i] - M1y.a[i]) / (2.0 * dy);
      comb[4].a[i] = (-M2y.a[i] - M1y.a[i]) / (2.0 * dy);
    }
    HMat Z(r, 6 * n);
    std::copy(S0.a.begin(), S0.a.end(), Z.a.begin());
    for (int d = 0; d < 5; ++d) {
      HMat t = titus::matmul(S0, comb[d]);
      std::copy(t.a.begin(), t.a.end(), Z.a.begin() + (size_t)(d + 1) * r * n);
    }
    return Z;
  }
  HMat range_w(const HMat& L, int ncols, const HMat* against, double scale) const {
    static const std::vector<int> all;
    return titus::range_basis(L, ncols, against, scale, rank_tol, true, all);
  }

  // ---- device range basis -------------------------------------------------
  // out (C x ncols) = orthonormal basis of range(Kw) minus span(Xa) (C x q),
  // completed from the interior candidates.  Kw (C x k) is overwritten.
  // scale < 0: relative to sigma_1 of the projected block.
  int range_basis_dev(double* Kw, int k, const double* Xa, int q, int ncols, double* out) {
    double scale = -1.0;
    if (Xa && q > 0) {
      tn(Kw, k, Kw, k, dev::OP_NONE, 0);
      std::vector<double> g0 = fetch(0, (size_t)k * k);
      double tr = 0.0;
      for (int i = 0; i < k; ++i) tr += g0[(size_t)i * k + i];
      scale = std::sqrt(tr);                              // |K|_F
      for (int pass = 0; pass < 2; ++pass) {
        if ((size_t)q * k > arena_n) throw TitusError(TITUS_E_ARG, "arena too small");
        timed("gram", 8.0 * C * (q + k), 2.0 * C * q * k,
              [&] { dev::launch_tn(st, g, Xa, C, q, Kw, C, k, dev::OP_NONE, partial, arena); });
        timed("gemm", 8.0 * C * (q + 2 * k), 2.0 * C * q * k,
              [&] { dev::launch_gemm(st, C, Xa, C, q, arena, k, -1.0, 1.0, Kw, C, nullptr); });
      }
    }
    // shifted CholeskyQR: Q1 = K R1^-1, Q2 = Q1 R2^-1, R3 = chol(Q2^T Q2)
    auto chol_shift = [&](const HMat& G, HMat& R) {
      double tr = 0.0;
      for (int i = 0; i < G.n; ++i) tr += G(i, i);
      double s = tr > 0.0 ? 1e-14 * tr : 1e-300;
      for (int attempt = 0; attempt < 12; ++attempt, s *= 10.0) {
        HMat Gs = G;
        for (int i = 0; i < G.n; ++i) Gs(i, i) += s;
        if (titus::cholesky_upper(Gs, R)) return;
      }
      throw TitusError(TITUS_E_RANK, "shifted Cholesky failed");
    };
    HMat G1 = tn_fetch(Kw, k, Kw, k);
    double tr1 = 0.0;
    for (int i = 0; i < k; ++i) tr1 += G1(i, i);
    if (!(std::isfinite(tr1))) throw TitusError(TITUS_E_NONFINITE, "non-finite spatial block");
    HMat R1, R2, R3;
    const int nc = std::min<int>((int)cand.size(), ncols + 64 + (Xa ? q : 0));
    std::vector<int> cidx(cand.begin(), cand.begin() + nc);
    HMat T, Bc;
    int p = 0;
    if (tr1 > 0.0) {
      chol_shift(G1, R1);
      gemm(Kw, k, titus::inv_upper(R1), 1.0, 0.0, Q1, nullptr);
      HMat G2 = tn_fetch(Q1, k, Q1, k);
      chol_shift(G2, R2);
      gemm(Q1, k, titus::inv_upper(R2), 1.0, 0.0, Q2, nullptr);
      // G3 and the candidate rows of Q2 (and of Xa) in one round trip
      tn(Q2, k, Q2, k, dev::OP_NONE, 0);
      const size_t off_q = (size_t)k * k, off_x = off_q + (size_t)nc * k;
      if (off_x + (size_t)nc * q > arena_n) throw TitusError(TITUS_E_ARG, "arena too small");
      dev::launch_gather_rows(st, Q2, C, k, cand_d, nc, arena + off_q);
      if (Xa && q > 0) dev::launch_gather_rows(st, Xa, C, q, cand_d, nc, arena + off_x);
      std::vector<double> h = fetch(0, off_x + (size_t)nc * (Xa ? q : 0));
      HMat G3 = as_mat(h, 0, k, k);
      if (!titus::cholesky_upper(G3, R3)) chol_shift(G3, R3);
      HMat R = titus::matmul(R3, titus::matmul(R2, R1));
      HMat P, Zr;
      std::vector<double> sv;
      titus::svd(R, P, sv, Zr);
      const double sc = scale < 0.0 ? (sv.empty() ? 0.0 : sv[0]) : scale;
      if (sc > 0.0)
        for (double x : sv) if (x > rank_tol * sc) ++p;
      p = std::min(p, ncols);
      T = titus::matmul(titus::inv_upper(R3), titus::cols(P, 0, p));     // k x p
      HMat Qc = as_mat(h, off_q, nc, k);
      HMat Uc = titus::matmul(Qc, T);                                      // nc x p
      Bc = (Xa && q > 0) ? titus::hcat(as_mat(h, off_x, nc, q), Uc) : Uc;
    } else {
      T = HMat(k, 0);
      if (Xa && q > 0) {
        dev::launch_gather_rows(st, Xa, C, q, cand_d, nc, arena);
        Bc = as_mat(fetch(0, (size_t)nc * q), 0, nc, q);
      } else {
        Bc = HMat(nc, 0);
      }
    }
    const int need = ncols - p;
    std::vector<int> chosen;
    HMat Rinv;
    if (need > 0 && !titus::canonical_completion(Bc, cidx, need, false, chosen, Rinv))
      throw TitusError(TITUS_E_RANK, "canonical completion ran out of candidates");
    // completion columns: E_ch Rinv - Xa (Xa[ch]^T Rinv) - U (U[ch]^T Rinv)
    std::vector<int> chpos;
    for (int c : chosen) chpos.push_back((int)(std::find(cidx.begin(), cidx.end(), c) - cidx.begin()));
    if (p > 0) {
      HMat Tfull(k, ncols);
      std::copy(T.a.begin(), T.a.end(), Tfull.a.begin());
      if (need > 0) {
        HMat Uch(need, p);
        const int off = (Xa && q > 0) ? q : 0;
        for (int s = 0; s < need; ++s)
          for (int j = 0; j < p; ++j) Uch(s, j) = Bc(chpos[s], off + j);
        HMat Tc = titus::matmul(T, titus::matmul(Uch, Rinv, true, false));   // k x need
        for (size_t i = 0; i < Tc.a.size(); ++i) Tfull.a[(size_t)p * k + i] = -Tc.a[i];
      }
      gemm(Q2, k, Tfull, 1.0, 0.0, out, nullptr);
    } else {
      // no range: start from zero columns (a zero-coefficient product keeps
      // the path free of memset nodes)
      gemm(Kw, k, HMat(k, ncols), 1.0, 0.0, out, nullptr);
    }
    if (need > 0) {
      if (Xa && q > 0) {
        HMat Xch(need, q);
        for (int s = 0; s < need; ++s)
          for (int j = 0; j < q; ++j) Xch(s, j) = Bc(chpos[s], j);
        HMat Tx = titus::matmul(Xch, Rinv, true, false);                     // q x need
        for (double& v : Tx.a) v = -v;
        gemm(Xa, q, Tx, 1.0, 1.0, out + (size_t)p * C, nullptr);
      }
      std::vector<int> rows, colsv;
      std::vector<double> vals;
      for (int s = 0; s < need; ++s)
        for (int t = s; t < need; ++t) {
          rows.push_back(chosen[s]); colsv.push_back(p + t); vals.push_back(Rinv(s, t));
        }
      const int ns = (int)rows.size();
      if ((size_t)ns > scat_n) throw TitusError(TITUS_E_ARG, "scatter area too small");
      HIPCHK(hipMemcpyAsync(srows, rows.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(scols, colsv.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(svals, vals.data(), ns * sizeof(double), hipMemcpyHostToDevice, st));
      dev::launch_scatter_add(st, out, C, srows, scols, svals, ns);
    }
    return p;
  }

  // ---- one energy step ------------------------------------------------------
  std::vector<double> sig(int s) const {
    return std::vector<double>(sigma.begin() + (size_t)s * nsig, sigma.begin() + (size_t)(s + 1) * nsig);
  }
  void dose_half(double Sn) {
    std::vector<double> v(rank, 0.0);
    for (int a = 0; a < rank; ++a)
      for (int b = 0; b < rank; ++b) v[a] += S(a, b) * W(0, b);
    const double* vd = upload_vec(v, 0);
    const double coef = 0.5 * dE * Sn;
    timed("dose", 8.0 * C * (rank + nq + 3), 2.0 * C * (rank + nq), [&] {
      dev::launch_dose(st, g, dose, X, C, rank, vd, psi, nq, m0, coef);
    });
  }
  void begin_step(int s, std::vector<double>& Dvec, std::vector<double>& Dout) {
    const std::vector<double> sg = sig(s);
    timed("bc", 16.0 * nb * nq + 4.0 * nb, 1.0 * nb * nq, [&] { dev::launch_bc(st, psi, C, nq, bcells, nb, bc_base, bc_scale[s - 1]); });
    dose_half(stopping[s - 1]);
    const double denom = 1.0 + dE * sg[0];
    timed("sweep", 16.0 * C * nq + 9.0 * C, 12.0 * C * nq, [&] {
      dev::launch_sweep(st, g, psi, psi2, nq, omega, dE, denom, restore);
    });
    std::swap(psi, psi2);
    Dvec.resize(nm);
    Dout.resize(nm);
    for (int i = 0; i < nm; ++i) {
      Dvec[i] = sg[degree[i]];
      Dout[i] = sg[0] - Dvec[i];
    }
  }
  void out_scatter(const std::vector<double>& Dout) {
    // L = W S^T, L[:,i] = (I + dE D) \ L[:,i] (SolverCSD.jl:1067-1074)
    HMat L = titus::matmul(W, S, false, true);
    for (int j = 0; j < L.n; ++j)
      for (int i = 0; i < nm; ++i) L(i, j) /= (1.0 + dE * Dout[i]);
    W = range_w(L, rank, nullptr, -1.0);
    S = titus::matmul(L, W, true, false);
  }
  HMat mtdw(const std::vector<double>& Dvec, const HMat& Wm) const {
    HMat DW = Wm;
    for (int j = 0; j < DW.n; ++j)
      for (int i = 0; i < nm; ++i) DW(i, j) *= Dvec[i];
    return titus::matmul(M, DW, true, false);                  // nq x r
  }
  // K = X S + dE psi MtDW with ghost rows zeroed (SolverCSD.jl:1079-1083)
  void scatter_K(const HMat& MtDW) {
    gemm(X, rank, S, 1.0, 0.0, K, nullptr);
    HMat T = MtDW;
    for (double& v : T.a) v *= dE;
    gemm(psi, nq, T, 1.0, 1.0, K, bmask, (size_t)rank * rank);
  }

  void step_fixed(int s) {
    std::vector<double> Dvec, Dout;
    begin_step(s, Dvec, Dout);
    const int r = rank;
    if (s > 1) {
      // K-step (SolverCSD.jl:1019-1031)
      std::vector<HMat> WAW = galerkin_w(W, W);
      const double* Zd = upload(stream_Z(S, WAW));
      timed("stream_K", 8.0 * C * (2 * r + 1), 12.0 * C * r * r,
            [&] { dev::launch_stream_k(st, g, X, C, r, Zd, r, dE, K, C); });
      range_basis_dev(K, r, nullptr, 0, r, Xn);
      HMat MUp = tn_fetch(Xn, r, X, r);
      // L-step (:1033-1046)
      std::vector<HMat> XLX = galerkin_x(X, r, X, r);
      HMat L = stream_L(titus::matmul(W, S, false, true), XLX);
      HMat WNew = range_w(L, r, nullptr, -1.0);
      HMat NUp = titus::matmul(WNew, W, true, false);
      W = WNew;
      std::swap(X, Xn);
      // S-step (:1051-1063)
      S = titus::matmul(titus::matmul(MUp, S), NUp, false, true);
      XLX = galerkin_x(X, r, X, r);
      WAW = galerkin_w(W, W);
      S = stream_S(S, XLX, WAW);
    }
    out_scatter(Dout);
    // in-scattering (:1079-1102)
    HMat MtDW = mtdw(Dvec, W);
    scatter_K(MtDW);
    range_basis_dev(K, r, nullptr, 0, r, Xn);
    HMat MUp = tn_fetch(Xn, r, X, r);
    HMat PtX = tn_fetch(psi, nq, X, r);                       // psi^T X, old X
    HMat L = titus::matmul(W, S, false, true);
    HMat MP = titus::matmul(M, PtX);
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < nm; ++i) L(i, j) += dE * Dvec[i] * MP(i, j);
    HMat WNew = range_w(L, r, nullptr, -1.0);
    HMat NUp = titus::matmul(WNew, W, true, false);
    W = WNew;
    std::swap(X, Xn);
    S = titus::matmul(titus::matmul(MUp, S), NUp, false, true);
    HMat XtP = tn_fetch(X, r, psi, nq);
    HMat add = titus::matmul(XtP, mtdw(Dvec, W));
    for (size_t i = 0; i < S.a.size(); ++i) S.a[i] += dE * add.a[i];
    dose_half(stopping[s]);
  }

  // truncate! (SolverCSD.jl:3369-3399) of the 2r x 2r Sup over the augmented
  // bases X1 (device, C x 2r) and W1; writes X (device), W, S, rank.
  void truncate_to(const HMat& Sup, const double* X1, const HMat& W1) {
    HMat U, V;
    std::vector<double> D;
    titus::svd_completed(Sup, rank_tol, U, D, V);
    const int n = (int)D.size();
    double nrm = 0.0;
    for (double d : D) nrm += d * d;
    nrm = std::sqrt(nrm);
    const double tol = theta * std::pow(nrm, theta_index);
    int rmax = n / 2;
    for (int j = 0; j < n; ++j) {
      double tail = 0.0;
      for (int i = j; i < n; ++i) tail += D[i];
      if (std::fabs(tail) < tol) { rmax = j + 1; break; }
    }
    rmax = std::max(std::min(rmax, rank_cap), 2);
    HMat Uk = titus::cols(U, 0, rmax);
    gemm(X1, n, Uk, 1.0, 0.0, X, nullptr);
    W = titus::matmul(W1, titus::cols(V, 0, rmax));
    S = HMat(rmax, rmax);
    for (int i = 0; i < rmax; ++i) S(i, i) = D[i];
    rank = rmax;
  }
  // X1 = [X | basis of range(K) minus span(X)] into Xn; returns nothing
  void augment_x() {
    HIPCHK(hipMemcpyAsync(Xn, X, (size_t)C * rank * sizeof(double), hipMemcpyDeviceToDevice, st));
    range_basis_dev(K, rank, X, rank, rank, Xn + (size_t)rank * C);
  }
  HMat augment_w(const HMat& L) const {
    const double Ln = titus::frob(L);
    return titus::hcat(W, range_w(L, rank, &W, Ln));
  }

  void step_bug(int s) {
    ranks.push_back(rank);
    std::vector<double> Dvec, Dout;
    begin_step(s, Dvec, Dout);
    if (s > 1) {
      const int r = rank;
      std::vector<HMat> XLX = galerkin_x(X, r, X, r);
      std::vector<HMat> WAW = galerkin_w(W, W);
      const double* Zd = upload(stream_Z(S, WAW));
      timed("stream_K", 8.0 * C * (2 * r + 1), 12.0 * C * r * r,
            [&] { dev::launch_stream_k(st, g, X, C, r, Zd, r, dE, K, C); });
      augment_x();                                            // Xn = X1 (C x 2r)
      HMat Mu = tn_fetch(Xn, 2 * r, X, r);
      HMat W1 = augment_w(stream_L(titus::matmul(W, S, false, true), XLX));
      HMat Nu = titus::matmul(W1, W, true, false);
      std::swap(X, Xn);                                       // X = X1
      W = W1;
      S = titus::matmul(titus::matmul(Mu, S), Nu, false, true);   // 2r x 2r
      XLX = galerkin_x(X, 2 * r, X, 2 * r);
      WAW = galerkin_w(W, W);
      HMat Sup = stream_S(S, XLX, WAW);
      // eta diagnostic (:2210-2220): rows r.. of X1^T L X1, columns r.. of W1^T A W1
      HMat acc(r, r);
      for (int k : {2, 3, 0, 1}) {
        HMat xl(r, 2 * r), wa(2 * r, r);
        for (int j = 0; j < 2 * r; ++j)
          for (int i = 0; i < r; ++i) xl(i, j) = XLX[k](r + i, j);
        for (int j = 0; j < r; ++j)
          for (int i = 0; i < 2 * r; ++i) wa(i, j) = WAW[k](i, r + j);
        HMat t = titus::matmul(titus::matmul(xl, S), wa);
        for (size_t i = 0; i < acc.a.size(); ++i) acc.a[i] += t.a[i];
      }
      etas.push_back(titus::frob(acc));
      // X currently holds X1; truncate writes the new basis into Xn, then swap
      std::swap(X, Xn);                                       // Xn = X1
      HMat W1c = W;
      truncate_to(Sup, Xn, W1c);
    }
    out_scatter(Dout);
    // in-scattering (:2252-2306)
    const int r = rank;
    HMat MtDW = mtdw(Dvec, W);
    HMat Xtpsi = tn_fetch(X, r, psi, nq);                    // old X
    scatter_K(MtDW);
    augment_x();
    HMat Mu = tn_fetch(Xn, 2 * r, X, r);
    HMat L = titus::matmul(W, S, false, true);
    HMat MX = titus::matmul(M, Xtpsi, false, true);
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < nm; ++i) L(i, j) += dE * Dvec[i] * MX(i, j);
    HMat W1 = augment_w(L);
    HMat Nu = titus::matmul(W1, W, true, false);
    S = titus::matmul(titus::matmul(Mu, S), Nu, false, true);
    HMat X1tpsi = tn_fetch(Xn, 2 * r, psi, nq);
    HMat add = titus::matmul(X1tpsi, mtdw(Dvec, W1));
    HMat Sup = S;
    for (size_t i = 0; i < Sup.a.size(); ++i) Sup.a[i] += dE * add.a[i];
    truncate_to(Sup, Xn, W1);
    dose_half(stopping[s]);
  }

  void step(int s) {
    if (s != steps_done + 1) throw TitusError(TITUS_E_STATE, "steps must run in order");
    if (s >= ne) throw TitusError(TITUS_E_STATE, "step beyond the energy grid");
    if (integrator == TITUS_INTEGRATOR_FIXED) step_fixed(s);
    else step_bug(s);
    steps_done = s;
    for (double v : S.a)
      if (!std::isfinite(v)) throw TitusError(TITUS_E_NONFINITE, "non-finite coefficient matrix");
  }
};

// ---------------------------------------------------------------------------
namespace {

template <class F>
int guarded(F&& f) {
  try {
    f();
    return TITUS_OK;
  } catch (const TitusError& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return TITUS_E_HIP;
  }
}

void require(bool ok, const char* what) {
  if (!ok) throw TitusError(TITUS_E_ARG, what);
}

HMat copy_mat(const double* p, int m, int n) {
  HMat A(m, n);
  if (p) std::memcpy(A.a.data(), p, sizeof(double) * (size_t)m * n);
  return A;
}

}  // namespace

extern "C" {

int32_t titus_abi_version(void) { return TITUS_ABI_VERSION; }

const char* titus_last_error(void) { return g_last_error.c_str(); }

int titus_create(titus_ctx** out, const titus_problem_v1* p, int32_t device) {
  return guarded([&] {
    require(out != nullptr && p != nullptr, "null argument");
    *out = nullptr;
    require(p->version == TITUS_ABI_VERSION, "problem struct version mismatch");
    require(p->nx >= 3 && p->ny >= 3, "need at least 3 cells per direction");
    require(p->dx > 0 && p->dy > 0, "cell widths must be positive");
    require(p->nm >= 1 && p->nq >= 1 && p->ne >= 3 && p->nsigma >= 1, "empty dimension");
    require(p->density && p->degree && p->Ax && p->Az && p->absAx && p->absAz && p->omega &&
                p->M && p->stopping && p->sigma && p->bcells && p->bc_base && p->bc_scale &&
                p->cand_cells, "missing operand");
    require(p->integrator == TITUS_INTEGRATOR_FIXED || p->integrator == TITUS_INTEGRATOR_BUG,
            "unknown integrator");
    require(p->precision == TITUS_PRECISION_FP64, "only the FP64 form is available in this build");
    require(p->r >= 1, "rank must be >= 1");
    const int C = p->nx * p->ny;
    for (int c = 0; c < C; ++c)
      require(std::isfinite(p->density[c]) && p->density[c] > 0.0, "density must be positive and finite");
    for (int i = 0; i < p->nm; ++i)
      require(p->degree[i] >= 0 && p->degree[i] < p->nsigma, "moment degree outside sigma table");
    for (int b = 0; b < p->nb; ++b) require(p->bcells[b] >= 0 && p->bcells[b] < C, "ghost cell out of range");
    for (int b = 0; b < p->nrestore; ++b)
      require(p->restore_cells[b] >= 0 && p->restore_cells[b] < C, "restore cell out of range");
    for (int b = 0; b < p->ncand; ++b) require(p->cand_cells[b] >= 0 && p->cand_cells[b] < C, "candidate out of range");

    std::unique_ptr<titus_ctx> ctx(new titus_ctx());
    titus_ctx& x = *ctx;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw TitusError(TITUS_E_NODEVICE, "no HIP device available");
    require(device >= 0 && device < ndev, "device index out of range");
    x.device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&x.st, hipStreamNonBlocking));
    x.nx = p->nx; x.ny = p->ny; x.C = C; x.nm = p->nm; x.nq = p->nq; x.ne = p->ne;
    x.nsig = p->nsigma; x.integrator = p->integrator; x.precision = p->precision;
    x.dx = p->dx; x.dy = p->dy; x.dE = p->dE; x.theta = p->theta; x.theta_index = p->theta_index;
    x.rank_tol = p->rank_tol > 0.0 ? p->rank_tol : 1e-7;
    x.rank = p->r;
    x.rank_cap = p->integrator == TITUS_INTEGRATOR_BUG ? p->rank_cap : p->r;
    require(x.rank_cap >= x.rank, "rank cap below the initial rank");
    x.capx = p->integrator == TITUS_INTEGRATOR_BUG ? 2 * x.rank_cap : x.rank;
    require(x.capx <= p->nm || p->integrator == TITUS_INTEGRATOR_FIXED, "2 * rank cap exceeds the moment count");
    require(x.rank <= p->nm, "rank exceeds the moment count");
    require(x.capx + 8 <= p->ncand, "too few completion candidates for the rank cap");
    x.stopping.assign(p->stopping, p->stopping + p->ne);
    x.sigma.assign(p->sigma, p->sigma + (size_t)p->ne * p->nsigma);
    x.bc_scale.assign(p->bc_scale, p->bc_scale + p->ne);
    x.degree.assign(p->degree, p->degree + p->nm);
    x.cand.assign(p->cand_cells, p->cand_cells + p->ncand);
    x.Ax = copy_mat(p->Ax, p->nm, p->nm);
    x.Az = copy_mat(p->Az, p->nm, p->nm);
    x.aAx = copy_mat(p->absAx, p->nm, p->nm);
    x.aAz = copy_mat(p->absAz, p->nm, p->nm);
    x.M = copy_mat(p->M, p->nm, p->nq);
    x.nb = p->nb;

    const size_t cx = (size_t)C * x.capx;
    x.rho = dalloc<double>(C);
    x.omega = dalloc<double>((size_t)p->nq * 3);
    x.bc_base = dalloc<double>((size_t)p->nb * p->nq);
    x.m0 = dalloc<double>(p->nq);
    x.psi = dalloc<double>((size_t)C * p->nq);
    x.psi2 = dalloc<double>((size_t)C * p->nq);
    x.dose = dalloc<double>(C);
    x.X = dalloc<double>(cx);
    x.Xn = dalloc<double>(cx);
    x.K = dalloc<double>(cx);
    x.Q1 = dalloc<double>(cx);
    x.Q2 = dalloc<double>(cx);
    const int wmax = std::max(x.capx, p->nq);
    x.partial = dalloc<double>(dev::tn_partial_size(C, wmax, wmax));
    const int ncand_used = std::min<int>(p->ncand, 2 * x.capx + 64);
    x.arena_n = std::max<size_t>(4 * (size_t)wmax * wmax, (size_t)x.capx * x.capx + (size_t)ncand_used * 2 * x.capx);
    x.arena = dalloc<double>(x.arena_n);
    x.upl_n = std::max<size_t>(6 * (size_t)x.capx * x.capx + 64, (size_t)x.capx * (x.capx + p->nq) + 2 * (size_t)wmax * wmax);
    x.upl = dalloc<double>(x.upl_n);
    x.scat_n = (size_t)x.capx * x.capx;
    x.srows = dalloc<int>(x.scat_n);
    x.scols = dalloc<int>(x.scat_n);
    x.svals = dalloc<double>(x.scat_n);
    x.bcells = dalloc<int>(p->nb);
    x.cand_d = dalloc<int>(p->ncand);
    x.bmask = dalloc<unsigned char>(C);
    x.restore = dalloc<unsigned char>(C);
    x.g = dev::Grid{p->nx, p->ny, C, p->dx, p->dy, x.rho};

    HIPCHK(hipMemcpy(x.rho, p->density, sizeof(double) * C, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.omega, p->omega, sizeof(double) * p->nq * 3, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.bc_base, p->bc_base, sizeof(double) * (size_t)p->nb * p->nq, hipMemcpyHostToDevice));
    std::vector<double> m0(p->nq);
    for (int q = 0; q < p->nq; ++q) m0[q] = p->M[(size_t)q * p->nm + 0];
    HIPCHK(hipMemcpy(x.m0, m0.data(), sizeof(double) * p->nq, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.bcells, p->bcells, sizeof(int) * p->nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.cand_d, p->cand_cells, sizeof(int) * p->ncand, hipMemcpyHostToDevice));
    std::vector<unsigned char> mask(C, 0), rest(C, 0);
    for (int b = 0; b < p->nb; ++b) mask[p->bcells[b]] = 1;
    for (int b = 0; b < p->nrestore; ++b) rest[p->restore_cells[b]] = 1;
    HIPCHK(hipMemcpy(x.bmask, mask.data(), C, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.restore, rest.data(), C, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(x.psi, 0, sizeof(double) * (size_t)C * p->nq));
    HIPCHK(hipMemset(x.psi2, 0, sizeof(double) * (size_t)C * p->nq));
    HIPCHK(hipMemset(x.dose, 0, sizeof(double) * C));
    HIPCHK(hipMemset(x.X, 0, sizeof(double) * cx));

    // initial factors (SolverCSD.jl:932-957): svd!(zeros) -> identity columns,
    // ghost rows of X zeroed
    const int r = x.rank;
    std::vector<double> X0((size_t)C * r, 0.0);
    if (p->X0) {
      std::memcpy(X0.data(), p->X0, sizeof(double) * X0.size());
    } else {
      for (int a = 0; a < r && a < C; ++a) X0[(size_t)a * C + a] = 1.0;
    }
    for (int b = 0; b < p->nb; ++b)
      for (int a = 0; a < r; ++a) X0[(size_t)a * C + p->bcells[b]] = 0.0;
    HIPCHK(hipMemcpy(x.X, X0.data(), sizeof(double) * X0.size(), hipMemcpyHostToDevice));
    x.S = p->S0 ? copy_mat(p->S0, r, r) : HMat(r, r);
    if (p->W0) {
      x.W = copy_mat(p->W0, p->nm, r);
    } else {
      x.W = HMat(p->nm, r);
      for (int a = 0; a < r; ++a) x.W(a, a) = 1.0;
    }
    HIPCHK(hipDeviceSynchronize());
    *out = ctx.release();
  });
}

int titus_step(titus_ctx* ctx, int32_t step) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    HIPCHK(hipSetDevice(ctx->device));
    ctx->step(step);
  });
}

int titus_solve(titus_ctx* ctx, int32_t last_step) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    HIPCHK(hipSetDevice(ctx->device));
    // the adaptive solver's loop breaks at n == nEnergies (SolverCSD.jl:2104)
    const int last_grid = ctx->integrator == TITUS_INTEGRATOR_BUG ? ctx->ne - 2 : ctx->ne - 1;
    const int last = last_step <= 0 ? last_grid : std::min<int>(last_step, ctx->ne - 1);
    for (int s = ctx->steps_done + 1; s <= last; ++s) ctx->step(s);
    HIPCHK(hipStreamSynchronize(ctx->st));
  });
}

int titus_progress(const titus_ctx* ctx, int32_t* steps_done, int32_t* rank) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    if (steps_done) *steps_done = ctx->steps_done;
    if (rank) *rank = ctx->rank;
  });
}

int titus_get_state(titus_ctx* ctx, double* dose, double* X, double* S, double* W, double* psi) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->st));
    const size_t C = ctx->C;
    if (dose) HIPCHK(hipMemcpy(dose, ctx->dose, sizeof(double) * C, hipMemcpyDeviceToHost));
    if (X) HIPCHK(hipMemcpy(X, ctx->X, sizeof(double) * C * ctx->rank, hipMemcpyDeviceToHost));
    if (psi) HIPCHK(hipMemcpy(psi, ctx->psi, sizeof(double) * C * ctx->nq, hipMemcpyDeviceToHost));
    if (S) std::memcpy(S, ctx->S.a.data(), sizeof(double) * ctx->S.a.size());
    if (W) std::memcpy(W, ctx->W.a.data(), sizeof(double) * ctx->W.a.size());
  });
}

int titus_get_history(const titus_ctx* ctx, int32_t* ranks, double* etas, int32_t n) {
  return guarded([&] {
    require(ctx != nullptr && n >= 0, "bad argument");
    for (int i = 0; i < n; ++i) {
      if (ranks) ranks[i] = i < (int)ctx->ranks.size() ? ctx->ranks[i] : 0;
      if (etas) etas[i] = i < (int)ctx->etas.size() ? ctx->etas[i] : 0.0;
    }
  });
}

int titus_set_timing(titus_ctx* ctx, int32_t enabled) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->timing.resolve();
    for (auto& t : ctx->timing.total_ms) t = 0.0;
    for (auto& l : ctx->timing.launches) l = 0;
    for (auto& b : ctx->timing.total_bytes) b = 0.0;
    for (auto& f : ctx->timing.total_flops) f = 0.0;
    ctx->timing.enabled = enabled != 0;
  });
}

int titus_kernel_stats(const titus_ctx* cctx, const char* name, double* total_ms,
                       int64_t* launches, double* bytes, double* flops) {
  return guarded([&] {
    require(cctx != nullptr && name != nullptr, "bad argument");
    titus_ctx* ctx = const_cast<titus_ctx*>(cctx);
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->timing.resolve();
    auto it = ctx->timing.ids.find(name);
    if (total_ms) *total_ms = it == ctx->timing.ids.end() ? 0.0 : ctx->timing.total_ms[it->second];
    const bool found = it != ctx->timing.ids.end();
    if (launches) *launches = found ? ctx->timing.launches[it->second] : 0;
    if (bytes) *bytes = found ? ctx->timing.total_bytes[it->second] : 0.0;
    if (flops) *flops = found ? ctx->timing.total_flops[it->second] : 0.0;
  });
}

void titus_destroy(titus_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->st) hipStreamSynchronize(ctx->st);
  delete ctx;
}

}  // extern "C"
