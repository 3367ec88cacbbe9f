// C-ABI and host orchestration of the DLRA energy step (see include/titus_hip.h).
//
// Division of work: every operation whose size grows with the cell count C
// (boundary source, dose, uncollided sweep, streaming stencils, Galerkin
// products, tall-skinny Gram and projection products, Cholesky passes, basis
// updates) runs on the device; the r x r, 2r x 2r and Nm x r algebra (angular
// factor, S updates, SVDs, completion) runs on the host in FP64.  Device
// results that the host needs are collected in one arena and fetched with a
// single transfer per phase: four host round trips per energy step.
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
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <atomic>
#include <thread>
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

template <class T>
T* halloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(hipHostMalloc(&p, n * sizeof(T), hipHostMallocDefault));
  return static_cast<T*>(p);
}

// Kernel timing with HIP events on the context's stream.
struct Timing {
  bool enabled = false;     // sampling this step
  int period = 0;           // time every period-th step (0: never)
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
  // host wall time of a phase (milliseconds), counted as one "launch"
  void host(const char* name, double ms) {
    const int k = id(name);
    total_ms[k] += ms;
    launches[k] += 1;
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

// Wall time of a host phase, recorded when the step is sampled.
struct HostScope {
  Timing& t;
  const char* name;
  std::chrono::steady_clock::time_point t0;
  HostScope(Timing& tm, const char* n) : t(tm), name(n), t0(std::chrono::steady_clock::now()) {}
  ~HostScope() {
    if (t.enabled)
      t.host(name, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
  }
};

// Device launches of one range basis, waiting for the host to finish it.
struct RangeJob {
  double* Kw = nullptr;
  int k = 0, q = 0, nc = 0;
  const double* Xa = nullptr;
  size_t g0 = 0, g1 = 0, r1 = 0, r2 = 0, r3 = 0, st = 0, rows_q = 0, rows_x = 0;
};

}  // namespace

struct titus_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  // problem (host copy)
  int nx = 0, ny = 0, C = 0, nm = 0, nq = 0, ne = 0, nsig = 0, integrator = 0, precision = 0;
  double dx = 0, dy = 0, dE = 0, theta = 0, theta_index = 0, rank_tol = 1e-7;
  int rank_cap = 0, capx = 0;
  std::vector<double> stopping, sigma, bc_scale;
  std::vector<int> degree, cand;
  HMat Ax, Az, aAx, aAz, M;
  // angular operators paired with the spatial stencils, in compressed rows:
  // |Ax|, |Az| (, |Ay|), Ax, Az (, Ay) -- dissipative first, then central
  std::vector<titus::SpMat> sA;
  int nd = 2;                 // spatial dimensions (3: k axis paired with Ay)
  int nz = 1;
  double dz = 1.0;
  int nops() const { return 2 * nd; }
  // term order of the streaming sums: central operators, then dissipative
  std::vector<int> term_order() const {
    std::vector<int> o;
    for (int d = 0; d < nd; ++d) o.push_back(nd + d);
    for (int d = 0; d < nd; ++d) o.push_back(d);
    return o;
  }
  // device buffers
  double *rho = nullptr, *irho = nullptr, *omega = nullptr, *bc_base = nullptr, *m0 = nullptr;
  void *psi = nullptr, *psi2 = nullptr;    // C x nq, FP64 (plain) or FP32 (mixed)
  dev::DT tpsi = dev::DT::F64;
  bool mixed = false;                       // psi stored and swept in FP32
  bool full32 = false;                      // streaming stencil, psi and Galerkin products in FP32
  double* dose = nullptr;
  double* u0p = nullptr;                    // l = 0 moment of psi, G partials x C
  int u0_groups = 1;
  dev::SweepPlan sweep_plan{};
  bool sweep_tuned = false;
  double *X = nullptr, *Xn = nullptr, *K = nullptr, *Q1 = nullptr, *Q2 = nullptr;
  double *partial = nullptr, *arena = nullptr, *upl = nullptr;
  int *bcells = nullptr, *cand_d = nullptr;
  unsigned char *bmask = nullptr, *restore = nullptr;
  // pinned host mirrors
  double *h_arena = nullptr, *h_upl = nullptr;
  size_t arena_n = 0, arena_top = 0, upl_n = 0, upl_top = 0;
  dev::Grid g{};
  int nb = 0;
  // state
  HMat S, W;
  int rank = 0, steps_done = 0;
  std::vector<int> ranks;
  std::vector<double> etas;
  mutable Timing timing;

  ~titus_ctx() {
    void* ptrs[] = {rho, irho, omega, bc_base, m0, psi, psi2, dose, u0p, X, Xn, K, Q1, Q2, partial, arena,
                    upl, bcells, cand_d, bmask, restore};
    for (void* p : ptrs) if (p) hipFree(p);
    void* hptrs[] = {h_arena, h_upl};
    for (void* p : hptrs) if (p) hipHostFree(p);
    if (st) hipStreamDestroy(st);
  }

  // ---- timing ---------------------------------------------------------------
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

  // ---- arena (device results for the host) and upload ring -------------------
  size_t aalloc(size_t n) {
    const size_t off = arena_top;
    arena_top += n;
    if (arena_top > arena_n) throw TitusError(TITUS_E_ARG, "result arena too small");
    return off;
  }
  void areset() { arena_top = 0; }
  // one transfer of everything the arena holds, then wait
  void fetch() {
    if (arena_top)
      HIPCHK(hipMemcpyAsync(h_arena, arena, arena_top * sizeof(double), hipMemcpyDeviceToHost, st));
    HostScope hs(timing, "host:wait");
    HIPCHK(hipStreamSynchronize(st));
    upl_top = 0;          // every queued upload has been consumed
  }
  HMat view(size_t off, int m, int n) const {
    HMat A(m, n);
    std::memcpy(A.a.data(), h_arena + off, sizeof(double) * (size_t)m * n);
    return A;
  }
  double* upload_raw(const double* src, size_t n) {
    if (upl_top + n > upl_n) {
      HIPCHK(hipStreamSynchronize(st));
      upl_top = 0;
      if (n > upl_n) throw TitusError(TITUS_E_ARG, "upload larger than the staging ring");
    }
    double* h = h_upl + upl_top;
    double* d = upl + upl_top;
    std::memcpy(h, src, n * sizeof(double));
    if (n) HIPCHK(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, st));
    upl_top += (n + 31) & ~size_t(31);
    return d;
  }
  double* upload(const HMat& h) { return upload_raw(h.a.data(), h.a.size()); }

  // ---- device building blocks --------------------------------------------------
  // arena[off : off + m*n] = A^T op(B)
  size_t tn(const double* A, int m, const double* B, int n, int op) {
    const int nops = op == dev::OP_ALL ? this->nops() : 1;
    const size_t off = aalloc((size_t)nops * m * n);
    const double by = 8.0 * C * (A == B ? m : m + n) + (op == dev::OP_NONE ? 0.0 : 8.0 * C);
    // Galerkin products run in FP32 only in the MIXED_FULL_FP32 form; Gram
    // and projection products of the QR passes are always FP64
    const bool fp32 = full32 && op != dev::OP_NONE;
    timed(op == dev::OP_NONE ? "gram" : "galerkin", by, 2.0 * nops * C * m * n, [&] {
      dev::launch_tn(st, g, A, dev::DT::F64, C, m, B, dev::DT::F64, C, n, op, fp32, partial,
                     arena + off);
    });
    return off;
  }
  // [Xa^T L_o Xb for the 2 nd stencil operators] at the returned offset
  size_t galerkin_x(const double* Xa, int ma, const double* Xb, int mb) {
    return tn(Xa, ma, Xb, mb, dev::OP_ALL);
  }
  std::vector<HMat> galerkin_view(size_t off, int ma, int mb) const {
    std::vector<HMat> out;
    for (int op = 0; op < nops(); ++op) out.push_back(view(off + (size_t)op * ma * mb, ma, mb));
    return out;
  }
  void gemm_dev(const double* A, int k, const double* Td, int n, double alpha, double beta,
                double* out, const unsigned char* mask) {
    timed("gemm", 8.0 * C * (k + n * (beta != 0.0 ? 2 : 1)), 2.0 * C * k * n, [&] {
      dev::launch_gemm(st, C, A, C, k, Td, n, alpha, beta, out, C, mask);
    });
  }
  void gemm(const double* A, int k, const HMat& T, double alpha, double beta, double* out,
            const unsigned char* mask) {
    gemm_dev(A, k, upload(T), T.n, alpha, beta, out, mask);
  }

  // ---- host building blocks (angular side) -------------------------------------
  // [A_o Wm for every angular operator]: computed once per basis and reused
  // by the Galerkin products, the L-step (A (W S^T) = (A W) S^T) and the
  // augmented basis (A [W Wt] = [A W, A Wt])
  using AWs = std::vector<HMat>;
  AWs apply_a(const HMat& Wm) const {
    HostScope hs(timing, "host:galerkin_w");
    AWs out;
    for (const titus::SpMat& A : sA) out.push_back(titus::spmm(A, Wm));
    return out;
  }
  std::vector<HMat> galerkin_from(const HMat& Wa, const AWs& awb) const {
    HostScope hs(timing, "host:galerkin_w");
    std::vector<HMat> out;
    for (const HMat& aw : awb) out.push_back(titus::matmul(Wa, aw, true));
    return out;
  }
  std::vector<HMat> galerkin_w(const HMat& Wa, const HMat& Wb) const {
    return galerkin_from(Wa, apply_a(Wb));
  }
  static AWs hcat_all(const AWs& a, const AWs& b) {
    AWs out;
    for (size_t o = 0; o < a.size(); ++o) out.push_back(titus::hcat(a[o], b[o]));
    return out;
  }
  // L - dE sum_o A_o L (X^T L_o X)^T (2D: Ax, Az, |Ax|, |Az| terms) for
  // L = W S^T, with A L = (A W) S^T from aw = apply_a(W)
  HMat stream_L(const AWs& aw, const std::vector<HMat>& XLX) const {
    HostScope hs(timing, "host:stream_LS");
    HMat out = titus::matmul(W, S, false, true);
    for (int k : term_order()) {
      HMat t = titus::matmul(titus::matmul(aw[k], S, false, true), XLX[k], false, true);
      for (size_t i = 0; i < out.a.size(); ++i) out.a[i] -= dE * t.a[i];
    }
    return out;
  }
  HMat stream_S(const HMat& S0, const std::vector<HMat>& XLX, const std::vector<HMat>& WAW) const {
    HostScope hs(timing, "host:stream_LS");
    HMat out = S0;
    for (int k : term_order()) {
      if (XLX[k].m != S0.m || XLX[k].n != S0.m || WAW[k].m != S0.n || WAW[k].n != S0.n)
        throw TitusError(TITUS_E_ARG, "S-step: Galerkin blocks do not match the coefficients");
      HMat t = titus::matmul(titus::matmul(XLX[k], S0), WAW[k]);
      for (size_t i = 0; i < out.a.size(); ++i) out.a[i] -= dE * t.a[i];
    }
    return out;
  }
  // Z = [S, S C_c, S C_x+, S C_x-, S C_y+, S C_y- (, S C_z+, S C_z-)] with
  // C_c = sum_d M1_d / h_d, C_d+ = (M2_d - M1_d) / (2 h_d) and
  // C_d- = (-M2_d - M1_d) / (2 h_d); M1_d = W^T |A_d| W, M2_d = W^T A_d W
  // (2D: C_c = M1x/dx + M1y/dy, then the east, west, north, south blocks)
  HMat stream_Z(const HMat& S0, const std::vector<HMat>& WAW) const {
    const int r = S0.m, n = S0.n;
    const double h[3] = {dx, dy, dz};
    const int nb = 1 + 2 * nd;
    std::vector<HMat> comb(nb, HMat(n, n));
    for (size_t i = 0; i < (size_t)n * n; ++i) {
      double c = WAW[0].a[i] / h[0];
      for (int d = 1; d < nd; ++d) c += WAW[d].a[i] / h[d];
      comb[0].a[i] = c;
      for (int d = 0; d < nd; ++d) {
        const double m1 = WAW[d].a[i], m2 = WAW[nd + d].a[i];
        comb[1 + 2 * d].a[i] = (m2 - m1) / (2.0 * h[d]);
        comb[2 + 2 * d].a[i] = (-m2 - m1) / (2.0 * h[d]);
      }
    }
    HMat Z(r, (nb + 1) * n);
    std::copy(S0.a.begin(), S0.a.end(), Z.a.begin());
    for (int d = 0; d < nb; ++d) {
      HMat t = titus::matmul(S0, comb[d]);
      std::copy(t.a.begin(), t.a.end(), Z.a.begin() + (size_t)(d + 1) * r * n);
    }
    return Z;
  }
  HMat range_w(const HMat& L, int ncols, const HMat* against, double scale) const {
    HostScope hs(timing, "host:range_w");
    static const std::vector<int> all;
    return titus::range_basis(L, ncols, against, scale, rank_tol, true, all);
  }
  void stream_k(int r, const std::vector<HMat>& WAW) {
    const double* Zd = upload(stream_Z(S, WAW));
    timed("stream_K", 8.0 * C * (2 * r + 1), 2.0 * (2 * nd + 2) * C * r * r,
          [&] { dev::launch_stream_k(st, g, X, C, r, Zd, r, dE, K, C, full32); });
  }

  // ---- device range basis --------------------------------------------------------
  // Launch the device part of the basis of range(Kw) minus span(Xa): the
  // projection (twice), three Gram + shifted-Cholesky passes and the candidate
  // rows; everything the host needs lands in the arena.  Kw is overwritten.
  RangeJob range_launch(double* Kw, int k, const double* Xa, int q) {
    RangeJob j;
    j.Kw = Kw; j.k = k; j.Xa = (Xa && q > 0) ? Xa : nullptr; j.q = j.Xa ? q : 0;
    if (j.Xa) {
      j.g0 = tn(Kw, k, Kw, k, dev::OP_NONE);              // |K|_F for the rank scale
      for (int pass = 0; pass < 2; ++pass) {
        const size_t p = tn(Xa, q, Kw, k, dev::OP_NONE);
        gemm_dev(Xa, q, arena + p, k, -1.0, 1.0, Kw, nullptr);
      }
    }
    const size_t sz = (size_t)k * k;
    j.g1 = tn(Kw, k, Kw, k, dev::OP_NONE);
    j.r1 = aalloc(sz); const size_t i1 = aalloc(sz);
    j.r2 = aalloc(sz); const size_t i2 = aalloc(sz);
    j.r3 = aalloc(sz); const size_t i3 = aalloc(sz);
    j.st = aalloc(4);
    timed("chol", 8.0 * 3 * sz, sz * k / 3.0, [&] {
      dev::launch_chol_inv(st, arena + j.g1, k, false, arena + j.r1, arena + i1, arena + j.st);
    });
    gemm_dev(Kw, k, arena + i1, k, 1.0, 0.0, Q1, nullptr);
    const size_t g2 = tn(Q1, k, Q1, k, dev::OP_NONE);
    timed("chol", 8.0 * 3 * sz, sz * k / 3.0, [&] {
      dev::launch_chol_inv(st, arena + g2, k, false, arena + j.r2, arena + i2, arena + j.st + 1);
    });
    gemm_dev(Q1, k, arena + i2, k, 1.0, 0.0, Q2, nullptr);
    const size_t g3 = tn(Q2, k, Q2, k, dev::OP_NONE);
    timed("chol", 8.0 * 3 * sz, sz * k / 3.0, [&] {
      dev::launch_chol_inv(st, arena + g3, k, true, arena + j.r3, arena + i3, arena + j.st + 2);
    });
    j.nc = std::min<int>((int)cand.size(), capx + 64 + j.q);
    j.rows_q = aalloc((size_t)j.nc * k);
    dev::launch_gather_rows(st, Q2, C, k, cand_d, j.nc, arena + j.rows_q);
    if (j.Xa) {
      j.rows_x = aalloc((size_t)j.nc * j.q);
      dev::launch_gather_rows(st, j.Xa, C, j.q, cand_d, j.nc, arena + j.rows_x);
    }
    return j;
  }

  // Host part after fetch(): SVD of R = R3 R2 R1, rank, canonical completion,
  // and the device launches that write the basis (C x ncols) to `out`.
  int range_finish(const RangeJob& j, int ncols, double* out) {
    HostScope hs(timing, "host:range_finish");
    const int k = j.k, q = j.q;
    HMat G1 = view(j.g1, k, k);
    double tr1 = 0.0;
    for (int i = 0; i < k; ++i) tr1 += G1(i, i);
    if (!std::isfinite(tr1)) throw TitusError(TITUS_E_NONFINITE, "non-finite spatial block");
    double scale = -1.0;
    if (j.Xa) {
      HMat G0 = view(j.g0, k, k);
      double t = 0.0;
      for (int i = 0; i < k; ++i) t += G0(i, i);
      scale = std::sqrt(t);
    }
    std::vector<int> cidx(cand.begin(), cand.begin() + j.nc);
    HMat T(k, 0), Bc;
    int p = 0;
    if (tr1 > 0.0) {
      for (int t = 0; t < 3; ++t)
        if (h_arena[j.st + t] < 0.0) throw TitusError(TITUS_E_RANK, "shifted Cholesky failed");
      HMat R1 = view(j.r1, k, k), R2 = view(j.r2, k, k), R3 = view(j.r3, k, k);
      HMat R = titus::matmul(R3, titus::matmul(R2, R1));
      HMat P, Zr;
      std::vector<double> sv;
      titus::svd(R, P, sv, Zr);
      const double sc = scale < 0.0 ? (sv.empty() ? 0.0 : sv[0]) : scale;
      if (sc > 0.0)
        for (double x : sv) if (x > rank_tol * sc) ++p;
      p = std::min(p, ncols);
      T = titus::matmul(titus::inv_upper(R3), titus::cols(P, 0, p));     // k x p
      HMat Uc = titus::matmul(view(j.rows_q, j.nc, k), T);               // nc x p
      Bc = j.Xa ? titus::hcat(view(j.rows_x, j.nc, q), Uc) : Uc;
    } else {
      Bc = j.Xa ? view(j.rows_x, j.nc, q) : HMat(j.nc, 0);
    }
    const int need = ncols - p;
    std::vector<int> chosen;
    HMat Rinv;
    if (need > 0 && !titus::canonical_completion(Bc, cidx, need, false, chosen, Rinv))
      throw TitusError(TITUS_E_RANK, "canonical completion ran out of candidates");
    std::vector<int> chpos;
    for (int c : chosen)
      chpos.push_back((int)(std::find(cidx.begin(), cidx.end(), c) - cidx.begin()));
    // range columns Q2 T; completion columns E_ch Rinv - Xa (Xa[ch]^T Rinv) - U (U[ch]^T Rinv)
    HMat Tfull(k, ncols);
    std::copy(T.a.begin(), T.a.end(), Tfull.a.begin());
    if (p > 0 && need > 0) {
      HMat Uch(need, p);
      for (int s = 0; s < need; ++s)
        for (int c = 0; c < p; ++c) Uch(s, c) = Bc(chpos[s], q + c);
      HMat Tc = titus::matmul(T, titus::matmul(Uch, Rinv, true, false));   // k x need
      for (size_t i = 0; i < Tc.a.size(); ++i) Tfull.a[(size_t)p * k + i] = -Tc.a[i];
    }
    // a zero block is written by a zero-coefficient product (no memset nodes)
    gemm((p > 0) ? Q2 : j.Kw, k, Tfull, 1.0, 0.0, out, nullptr);
    if (need > 0) {
      if (j.Xa) {
        HMat Xch(need, q);
        for (int s = 0; s < need; ++s)
          for (int c = 0; c < q; ++c) Xch(s, c) = Bc(chpos[s], c);
        HMat Tx = titus::matmul(Xch, Rinv, true, false);                     // q x need
        for (double& v : Tx.a) v = -v;
        gemm(j.Xa, q, Tx, 1.0, 1.0, out + (size_t)p * C, nullptr);
      }
      // (row, column, value) triples of the upper-triangular Rinv at the
      // chosen cells, packed as doubles through the upload ring
      const int ns = need * (need + 1) / 2;
      std::vector<double> pk(3 * (size_t)ns);
      int e = 0;
      for (int s = 0; s < need; ++s)
        for (int t = s; t < need; ++t, ++e) {
          pk[e] = chosen[s]; pk[ns + e] = p + t; pk[2 * ns + e] = Rinv(s, t);
        }
      dev::launch_scatter_add(st, out, C, upload_raw(pk.data(), pk.size()), ns);
    }
    return p;
  }

  // ---- pieces of one energy step ---------------------------------------------
  std::vector<double> sig(int s) const {
    return std::vector<double>(sigma.begin() + (size_t)s * nsig, sigma.begin() + (size_t)(s + 1) * nsig);
  }
  void dose_half(double Sn) {
    std::vector<double> v(rank, 0.0);
    for (int a = 0; a < rank; ++a)
      for (int b = 0; b < rank; ++b) v[a] += S(a, b) * W(0, b);
    const double* vd = upload_raw(v.data(), v.size());
    const double coef = 0.5 * dE * Sn;
    timed("dose", 8.0 * C * (rank + 3 + u0_groups), 2.0 * C * (rank + u0_groups), [&] {
      dev::launch_dose(st, g, dose, X, C, rank, vd, u0p, u0_groups, coef);
    });
  }
  void begin_step(int s, std::vector<double>& Dvec, std::vector<double>& Dout) {
    const std::vector<double> sg = sig(s);
    const double ps = mixed ? 4.0 : 8.0;
    timed("bc", (8.0 + ps) * nb * nq + 4.0 * nb + 8.0 * nb * u0_groups, 3.0 * nb * nq, [&] {
      dev::launch_bc(st, psi, tpsi, C, nq, bcells, nb, bc_base, bc_scale[s - 1], m0, u0p,
                     u0_groups);
    });
    dose_half(stopping[s - 1]);
    const double denom = 1.0 + dE * sg[0];
    // psi read and written once, density, restore flags and moment partials
    timed("sweep", 2.0 * ps * C * nq + 9.0 * C + 8.0 * C * u0_groups, 14.0 * C * nq, [&] {
      dev::launch_sweep(st, g, psi, psi2, tpsi, nq, omega, dE, denom, restore, m0, u0p, sweep_plan);
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
    const double* Td = upload(T);
    const double ps = mixed ? 4.0 : 8.0;
    timed("gemm", ps * C * nq + 16.0 * C * T.n, 2.0 * C * nq * T.n, [&] {
      dev::launch_gemm(st, C, psi, tpsi, C, nq, Td, T.n, 1.0, 1.0, K, C, bmask, full32);
    });
  }
  // arena[off : off + m*nq] = Xa^T psi (the psi contractions; FP32 in MIXED_FULL_FP32)
  size_t tn_psi(const double* Xa, int m) {
    const size_t off = aalloc((size_t)m * nq);
    const double ps = mixed ? 4.0 : 8.0;
    timed("gram", 8.0 * C * m + ps * C * nq, 2.0 * C * m * nq, [&] {
      dev::launch_tn(st, g, Xa, dev::DT::F64, C, m, psi, tpsi, C, nq, dev::OP_NONE, full32, partial,
                     arena + off);
    });
    return off;
  }

  void step_fixed(int s) {
    std::vector<double> Dvec, Dout;
    begin_step(s, Dvec, Dout);
    const int r = rank;
    if (s > 1) {
      // K-step and old Galerkin products (SolverCSD.jl:1019-1038)
      areset();
      const size_t gx = galerkin_x(X, r, X, r);
      const AWs aw = apply_a(W);
      stream_k(r, galerkin_from(W, aw));
      RangeJob j = range_launch(K, r, nullptr, 0);
      fetch();
      std::vector<HMat> XLX = galerkin_view(gx, r, r);
      range_finish(j, r, Xn);
      // L-step on the host while the device writes XNew (:1033-1046)
      HMat L = stream_L(aw, XLX);
      HMat WNew = range_w(L, r, nullptr, -1.0);
      HMat NUp = titus::matmul(WNew, W, true, false);
      areset();
      const size_t mu = tn(Xn, r, X, r, dev::OP_NONE);
      const size_t gn = galerkin_x(Xn, r, Xn, r);
      fetch();
      HMat MUp = view(mu, r, r);
      XLX = galerkin_view(gn, r, r);
      W = WNew;
      std::swap(X, Xn);
      // S-step (:1051-1063)
      S = titus::matmul(titus::matmul(MUp, S), NUp, false, true);
      S = stream_S(S, XLX, galerkin_w(W, W));
    }
    out_scatter(Dout);
    // in-scattering (:1079-1102)
    areset();
    const size_t xtp0 = tn_psi(X, r);                       // X^T psi, old X
    scatter_K(mtdw(Dvec, W));
    RangeJob j = range_launch(K, r, nullptr, 0);
    fetch();
    HMat PtX = titus::transpose(view(xtp0, r, nq));
    range_finish(j, r, Xn);
    HMat L = titus::matmul(W, S, false, true);
    HMat MP = titus::matmul(M, PtX);
    for (int c = 0; c < r; ++c)
      for (int i = 0; i < nm; ++i) L(i, c) += dE * Dvec[i] * MP(i, c);
    HMat WNew = range_w(L, r, nullptr, -1.0);
    HMat NUp = titus::matmul(WNew, W, true, false);
    areset();
    const size_t mu = tn(Xn, r, X, r, dev::OP_NONE);
    const size_t xtp = tn_psi(Xn, r);
    fetch();
    HMat MUp = view(mu, r, r);
    HMat XtP = view(xtp, r, nq);
    W = WNew;
    std::swap(X, Xn);
    S = titus::matmul(titus::matmul(MUp, S), NUp, false, true);
    HMat add = titus::matmul(XtP, mtdw(Dvec, W));
    for (size_t i = 0; i < S.a.size(); ++i) S.a[i] += dE * add.a[i];
    dose_half(stopping[s]);
  }

  // truncate! (SolverCSD.jl:3369-3399) of the 2r x 2r Sup over the augmented
  // bases X1 (device, C x 2r) and W1; writes X (device), W, S, rank.
  void truncate_to(const HMat& Sup, const double* X1, const HMat& W1) {
    HostScope hs(timing, "host:truncate");
    HMat U, V;
    std::vector<double> D;
    titus::svd_completed(Sup, rank_tol, U, D, V);
    const int n = (int)D.size();
    double nrm = 0.0;
    for (double d : D) nrm += d * d;
    nrm = std::sqrt(nrm);
    const double tol = theta * std::pow(nrm, theta_index);
    int rmax = n / 2;
    for (int jj = 0; jj < n; ++jj) {
      double tail = 0.0;
      for (int i = jj; i < n; ++i) tail += D[i];
      if (std::fabs(tail) < tol) { rmax = jj + 1; break; }
    }
    rmax = std::max(std::min(rmax, rank_cap), 2);
    gemm(X1, n, titus::cols(U, 0, rmax), 1.0, 0.0, X, nullptr);
    W = titus::matmul(W1, titus::cols(V, 0, rmax));
    S = HMat(rmax, rmax);
    for (int i = 0; i < rmax; ++i) S(i, i) = D[i];
    rank = rmax;
  }
  HMat augment_w(const HMat& L) const {
    return titus::hcat(W, range_w(L, rank, &W, titus::frob(L)));
  }
  void copy_x_to_aug(int r) {
    HIPCHK(hipMemcpyAsync(Xn, X, (size_t)C * r * sizeof(double), hipMemcpyDeviceToDevice, st));
  }

  void step_bug(int s) {
    ranks.push_back(rank);
    std::vector<double> Dvec, Dout;
    begin_step(s, Dvec, Dout);
    if (s > 1) {
      const int r = rank;
      areset();
      const size_t gx = galerkin_x(X, r, X, r);
      const AWs aw = apply_a(W);
      stream_k(r, galerkin_from(W, aw));
      copy_x_to_aug(r);
      RangeJob j = range_launch(K, r, X, r);
      fetch();
      std::vector<HMat> XLX = galerkin_view(gx, r, r);
      range_finish(j, r, Xn + (size_t)r * C);               // Xn = X1 = [X | Xt]
      HMat W1 = augment_w(stream_L(aw, XLX));
      HMat Nu = titus::matmul(W1, W, true, false);
      areset();
      const size_t mu = tn(Xn, 2 * r, X, r, dev::OP_NONE);
      const size_t g1 = galerkin_x(Xn, 2 * r, Xn, 2 * r);
      fetch();
      HMat Mu = view(mu, 2 * r, r);
      XLX = galerkin_view(g1, 2 * r, 2 * r);
      HMat S1 = titus::matmul(titus::matmul(Mu, S), Nu, false, true);   // 2r x 2r
      // A W1 = [A W, A Wt]
      std::vector<HMat> WAW =
          galerkin_from(W1, hcat_all(aw, apply_a(titus::cols(W1, r, W1.n))));
      HMat Sup = stream_S(S1, XLX, WAW);
      // eta diagnostic (:2210-2220): rows r.. of X1^T L X1, columns r.. of W1^T A W1
      HMat acc(r, r);
      for (int k : term_order()) {
        HMat xl(r, 2 * r), wa(2 * r, r);
        for (int c = 0; c < 2 * r; ++c)
          for (int i = 0; i < r; ++i) xl(i, c) = XLX[k](r + i, c);
        for (int c = 0; c < r; ++c)
          for (int i = 0; i < 2 * r; ++i) wa(i, c) = WAW[k](i, r + c);
        HMat t = titus::matmul(titus::matmul(xl, S1), wa);
        for (size_t i = 0; i < acc.a.size(); ++i) acc.a[i] += t.a[i];
      }
      etas.push_back(titus::frob(acc));
      truncate_to(Sup, Xn, W1);
    }
    out_scatter(Dout);
    // in-scattering (:2252-2306)
    const int r = rank;
    areset();
    const size_t xtp = tn_psi(X, r);                          // old X
    scatter_K(mtdw(Dvec, W));
    copy_x_to_aug(r);
    RangeJob j = range_launch(K, r, X, r);
    fetch();
    HMat Xtpsi = view(xtp, r, nq);
    range_finish(j, r, Xn + (size_t)r * C);
    HMat L = titus::matmul(W, S, false, true);
    HMat MX = titus::matmul(M, Xtpsi, false, true);
    for (int c = 0; c < r; ++c)
      for (int i = 0; i < nm; ++i) L(i, c) += dE * Dvec[i] * MX(i, c);
    HMat W1 = augment_w(L);
    HMat Nu = titus::matmul(W1, W, true, false);
    areset();
    const size_t mu = tn(Xn, 2 * r, X, r, dev::OP_NONE);
    const size_t x1p = tn_psi(Xn, 2 * r);
    fetch();
    HMat Mu = view(mu, 2 * r, r);
    HMat S1 = titus::matmul(titus::matmul(Mu, S), Nu, false, true);
    HMat add = titus::matmul(view(x1p, 2 * r, nq), mtdw(Dvec, W1));
    for (size_t i = 0; i < S1.a.size(); ++i) S1.a[i] += dE * add.a[i];
    truncate_to(S1, Xn, W1);
    dose_half(stopping[s]);
  }

  void step(int s) {
    if (s != steps_done + 1) throw TitusError(TITUS_E_STATE, "steps must run in order");
    if (s >= ne) throw TitusError(TITUS_E_STATE, "step beyond the energy grid");
    timing.enabled = timing.period > 0 && s % timing.period == 0;
    HostScope hs(timing, "host:step");
    if (integrator == TITUS_INTEGRATOR_FIXED) step_fixed(s);
    else step_bug(s);
    steps_done = s;
    for (double v : S.a)
      if (!std::isfinite(v)) throw TitusError(TITUS_E_NONFINITE, "non-finite coefficient matrix");
  }
};

// ---------------------------------------------------------------------------
// Launch-plan autotuner of the uncollided sweep (the kernel that moves the
// most bytes).  Each candidate plan (rows per strip x ordinate groups) is
// timed with HIP events on the context's own, still zero, psi buffers; the
// fastest is kept for every context of the same grid, ordinate count and
// precision in this process.  Problems below kTuneMinWork cells x ordinates
// are launch-latency bound and keep the heuristic plan.
namespace {
constexpr size_t kTuneMinWork = size_t(1) << 18;
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int>, dev::SweepPlan> g_tuned;
}  // namespace

static void tune_sweep(titus_ctx& x) {
  x.sweep_plan = dev::sweep_default(x.g, x.nq, x.tpsi);
  x.sweep_tuned = false;
  if ((size_t)x.C * x.nq < kTuneMinWork) return;
  const auto key = std::make_tuple(x.nx, x.ny * x.nz + (x.nz > 1 ? 1 << 24 : 0), x.nq, x.mixed ? 1 : 0);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) {
    x.sweep_plan = it->second;
    x.sweep_tuned = true;
    return;
  }
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  const double denom = 1.0 + x.dE * x.sigma[0];
  float best = 0.f;
  for (const dev::SweepPlan& c : dev::sweep_candidates(x.g, x.nq, x.tpsi)) {
    auto run = [&] {
      dev::launch_sweep(x.st, x.g, x.psi, x.psi2, x.tpsi, x.nq, x.omega, x.dE, denom, x.restore,
                        x.m0, x.u0p, c);
    };
    run();
    HIPCHK(hipEventRecord(a, x.st));
    for (int rep = 0; rep < 3; ++rep) run();
    HIPCHK(hipEventRecord(b, x.st));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    if (!x.sweep_tuned || ms < best) {
      best = ms;
      x.sweep_plan = c;
      x.sweep_tuned = true;
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  g_tuned[key] = x.sweep_plan;
}

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
    require(p->version == 1 || p->version == TITUS_ABI_VERSION, "problem struct version mismatch");
    const int nz = p->version >= 2 ? p->nz : 1;
    require(p->nx >= 3 && p->ny >= 3 && (nz == 1 || nz >= 3), "need at least 3 cells per direction");
    if (nz > 1) require(p->dz > 0 && p->Ay && p->absAy, "3D grid needs dz, Ay and |Ay|");
    require(p->dx > 0 && p->dy > 0, "cell widths must be positive");
    require(p->nm >= 1 && p->nq >= 1 && p->ne >= 3 && p->nsigma >= 1, "empty dimension");
    require(p->density && p->degree && p->Ax && p->Az && p->absAx && p->absAz && p->omega &&
                p->M && p->stopping && p->sigma && p->bcells && p->bc_base && p->bc_scale &&
                p->cand_cells, "missing operand");
    require(p->integrator == TITUS_INTEGRATOR_FIXED || p->integrator == TITUS_INTEGRATOR_BUG,
            "unknown integrator");
    require(p->precision == TITUS_PRECISION_FP64 || p->precision == TITUS_PRECISION_MIXED ||
                p->precision == TITUS_PRECISION_MIXED_FULL_FP32,
            "unknown precision form");
    require(p->r >= 1, "rank must be >= 1");
    require((long long)p->nx * p->ny * nz < (1LL << 31), "grid too large");
    const int C = p->nx * p->ny * nz;
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
    x.mixed = p->precision != TITUS_PRECISION_FP64;
    x.full32 = p->precision == TITUS_PRECISION_MIXED_FULL_FP32;
    x.tpsi = x.mixed ? dev::DT::F32 : dev::DT::F64;
    x.dx = p->dx; x.dy = p->dy; x.dE = p->dE; x.theta = p->theta; x.theta_index = p->theta_index;
    x.rank_tol = p->rank_tol > 0.0 ? p->rank_tol : 1e-7;
    x.rank = p->r;
    x.rank_cap = p->integrator == TITUS_INTEGRATOR_BUG ? p->rank_cap : p->r;
    require(x.rank_cap >= x.rank, "rank cap below the initial rank");
    x.capx = p->integrator == TITUS_INTEGRATOR_BUG ? 2 * x.rank_cap : x.rank;
    require(x.capx <= p->nm || p->integrator == TITUS_INTEGRATOR_FIXED, "2 * rank cap exceeds the moment count");
    require(x.rank <= p->nm, "rank exceeds the moment count");
    require(x.capx <= 96, "rank cap too large for the device Cholesky (2 r <= 96)");
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
    x.nz = nz;
    x.nd = nz > 1 ? 3 : 2;
    x.dz = nz > 1 ? p->dz : 1.0;
    if (x.nd == 3) {
      const HMat Ay = copy_mat(p->Ay, p->nm, p->nm), aAy = copy_mat(p->absAy, p->nm, p->nm);
      x.sA = {titus::SpMat::from_dense(x.aAx), titus::SpMat::from_dense(x.aAz),
              titus::SpMat::from_dense(aAy), titus::SpMat::from_dense(x.Ax),
              titus::SpMat::from_dense(x.Az), titus::SpMat::from_dense(Ay)};
    } else {
      x.sA = {titus::SpMat::from_dense(x.aAx), titus::SpMat::from_dense(x.aAz),
              titus::SpMat::from_dense(x.Ax), titus::SpMat::from_dense(x.Az)};
    }
    x.nb = p->nb;

    const size_t cx = (size_t)C * x.capx;
    x.rho = dalloc<double>(C);
    x.irho = dalloc<double>(C);
    x.omega = dalloc<double>((size_t)p->nq * 3);
    x.bc_base = dalloc<double>((size_t)p->nb * p->nq);
    x.m0 = dalloc<double>(p->nq);
    const size_t psz = x.mixed ? sizeof(float) : sizeof(double);
    x.psi = dalloc<unsigned char>((size_t)C * p->nq * psz);
    x.psi2 = dalloc<unsigned char>((size_t)C * p->nq * psz);
    x.dose = dalloc<double>(C);
    x.u0p = dalloc<double>((size_t)dev::kMaxSweepGroups * C);
    x.X = dalloc<double>(cx);
    x.Xn = dalloc<double>(cx);
    x.K = dalloc<double>(cx);
    x.Q1 = dalloc<double>(cx);
    x.Q2 = dalloc<double>(cx);
    const int wmax = std::max(x.capx, p->nq);
    x.partial = dalloc<double>(dev::tn_partial_size());
    const size_t w2 = (size_t)wmax * wmax;
    const size_t ncand_used = std::min<size_t>(p->ncand, (size_t)2 * x.capx + 64);
    // largest phase: Galerkin of the augmented basis plus one range job
    x.arena_n = 8 * w2 + 10 * (size_t)x.capx * x.capx + ncand_used * 2 * x.capx + 256;
    x.arena = dalloc<double>(x.arena_n);
    x.h_arena = halloc<double>(x.arena_n);
    x.upl_n = 64 * w2 + 8192;
    x.upl = dalloc<double>(x.upl_n);
    x.h_upl = halloc<double>(x.upl_n);
    x.bcells = dalloc<int>(p->nb);
    x.cand_d = dalloc<int>(p->ncand);
    x.bmask = dalloc<unsigned char>(C);
    x.restore = dalloc<unsigned char>(C);
    x.g = dev::Grid{p->nx, p->ny, C, p->dx, p->dy, x.rho, x.irho, x.nz, x.dz};

    HIPCHK(hipMemcpy(x.rho, p->density, sizeof(double) * C, hipMemcpyHostToDevice));
    {
      std::vector<double> ir(C);
      for (int c = 0; c < C; ++c) ir[c] = 1.0 / p->density[c];
      HIPCHK(hipMemcpy(x.irho, ir.data(), sizeof(double) * C, hipMemcpyHostToDevice));
    }
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
    HIPCHK(hipMemset(x.psi, 0, psz * (size_t)C * p->nq));
    HIPCHK(hipMemset(x.psi2, 0, psz * (size_t)C * p->nq));
    HIPCHK(hipMemset(x.dose, 0, sizeof(double) * C));
    // psi starts at zero, so does its moment
    HIPCHK(hipMemset(x.u0p, 0, sizeof(double) * dev::kMaxSweepGroups * C));
    tune_sweep(x);
    x.u0_groups = dev::sweep_groups(x.g, p->nq, x.sweep_plan);
    HIPCHK(hipMemset(x.u0p, 0, sizeof(double) * dev::kMaxSweepGroups * C));
    HIPCHK(hipMemset(x.psi2, 0, psz * (size_t)C * p->nq));
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

int titus_solve_batch(titus_ctx** ctxs, int32_t n, int32_t nthreads, int32_t* status) {
  if (!ctxs || n < 0) {
    g_last_error = "bad argument";
    return TITUS_E_ARG;
  }
  int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = std::max(1, std::min<int>(nt, n));
  std::vector<int> codes(n, TITUS_OK);
  std::vector<std::string> msgs(n);
  std::atomic<int> next{0};
  auto worker = [&] {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n) return;
      codes[i] = titus_solve(ctxs[i], 0);
      if (codes[i] != TITUS_OK) msgs[i] = g_last_error;   // thread-local message
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  int first = TITUS_OK;
  for (int i = 0; i < n; ++i) {
    if (status) status[i] = codes[i];
    if (first == TITUS_OK && codes[i] != TITUS_OK) {
      first = codes[i];
      g_last_error = "solve " + std::to_string(i) + ": " + msgs[i];
    }
  }
  return first;
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
    if (psi) {
      if (ctx->mixed) {
        std::vector<float> f(C * ctx->nq);
        HIPCHK(hipMemcpy(f.data(), ctx->psi, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < f.size(); ++i) psi[i] = f[i];
      } else {
        HIPCHK(hipMemcpy(psi, ctx->psi, sizeof(double) * C * ctx->nq, hipMemcpyDeviceToHost));
      }
    }
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

int titus_set_timing(titus_ctx* ctx, int32_t period) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    require(period >= 0, "timing period must be >= 0");
    HIPCHK(hipStreamSynchronize(ctx->st));
    ctx->timing.resolve();
    for (auto& t : ctx->timing.total_ms) t = 0.0;
    for (auto& l : ctx->timing.launches) l = 0;
    for (auto& b : ctx->timing.total_bytes) b = 0.0;
    for (auto& f : ctx->timing.total_flops) f = 0.0;
    ctx->timing.period = period;
    ctx->timing.enabled = false;
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
    const bool found = it != ctx->timing.ids.end();
    if (total_ms) *total_ms = found ? ctx->timing.total_ms[it->second] : 0.0;
    if (launches) *launches = found ? ctx->timing.launches[it->second] : 0;
    if (bytes) *bytes = found ? ctx->timing.total_bytes[it->second] : 0.0;
    if (flops) *flops = found ? ctx->timing.total_flops[it->second] : 0.0;
  });
}

int titus_get_tuning(const titus_ctx* ctx, const char* kernel, int32_t* values, int32_t n) {
  return guarded([&] {
    require(ctx != nullptr && kernel != nullptr && (values != nullptr || n == 0), "bad argument");
    require(std::string(kernel) == "sweep", "unknown kernel (tuned kernels: sweep)");
    const int32_t v[4] = {ctx->sweep_plan.rows, ctx->sweep_plan.groups, ctx->sweep_tuned ? 1 : 0,
                          ctx->sweep_plan.nontemporal ? 1 : 0};
    for (int i = 0; i < n && i < 4; ++i) values[i] = v[i];
  });
}

void titus_destroy(titus_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->st) hipStreamSynchronize(ctx->st);
  delete ctx;
}

}  // extern "C"
