// Host context of one DLRA solve: problem copy, device buffers and the
// energy-step orchestration shared by the forward solver (solver.cpp) and the
// adjoint (adjoint.cpp).
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
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
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
#include "angular.h"
#include "batch.h"
#include "kernels.h"
#include "smallla.h"

using titus::HMat;
namespace dev = titus::dev;

namespace titus_host {

inline thread_local std::string g_last_error;

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

// True when a context appears more than once in a batch (two host threads
// would then drive one context and its stream).
inline bool batch_has_duplicates(titus_ctx* const* ctxs, int n) {
  std::vector<const void*> v(ctxs, ctxs + n);
  std::sort(v.begin(), v.end());
  for (int i = 1; i < n; ++i)
    if (v[i] && v[i] == v[i - 1]) return true;
  return false;
}

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
    for (auto& r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
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

}  // namespace titus_host
using namespace titus_host;

// What the adjoint needs of one spatial basis update X1 = range_basis(K):
// the device CholeskyQR3 K = Q3 R with Q3 = Q2 R3^-1, the SVD R = P diag(sv)
// V^T (so U = Q3 P), the kept rank p and the coefficients Tp (p x ncols) of
// X1 on U[:, :p].  Q2 is a device copy (C x k) owned by the adjoint store.
struct SpatialOrthTape {
  int p = 0, k = 0;
  std::vector<double> sv;
  HMat P, V, R3inv, Tp;
  double* Q2 = nullptr;
};

// Intermediates of one fixed-rank energy step, recorded by step_fixed when
// the context's tape pointer is set (the adjoint's recomputation).  Device
// buffers (C x r and C x nq) belong to the adjoint store.
struct StepTape {
  int s = 0;
  double *X = nullptr, *X1 = nullptr, *X3 = nullptr, *psi_a = nullptr, *psi_b = nullptr;
  HMat S, W;                                   // at the start of the step
  std::vector<double> Dvec, Dout;
  double sig0 = 0.0;
  // streaming (s > 1)
  std::vector<HMat> WAW, XLX, XLX1, WAW1;
  HMat MUp1, NUp1, S1, L0, W1;
  titus::OrthRecord Rw1;
  SpatialOrthTape Rx1;
  // out- and in-scattering
  HMat Ws, Ss, Lo, W2, S3, MtDW, MUp3, W3, NUp3, B3, A3, S5;
  titus::OrthRecord Rw2, Rw3;
  SpatialOrthTape Rx3;
};

struct AdjointStore;
struct titus_ctx;
struct LockstepGroup;
// solver.cpp: the calling context's sync point in its lockstep group (flushes
// the recorded launches of every context of the group once all have arrived)
void lockstep_sync(titus_ctx& x);
// solver.cpp: run work(ctx, i) for every context of a lockstep batch (the
// validated contexts of titus_solve_batch_lockstep's contract): up to
// `width` at once in `groups` groups, their launches shared; per-context
// codes in status, counters in stats (6), tuned tiles in tiles (4 per group)
int lockstep_run(titus_ctx** ctxs, int n, int width, int groups, int32_t* status, int64_t* stats,
                 int32_t* tiles, const std::function<void(titus_ctx&, int)>& work);
// adjoint.cpp: for a context with checkpoints, the length of the last
// checkpoint segment of its solve so far (the first one its adjoint
// recomputes); 0 without checkpoints
int adjoint_phase_key(const titus_ctx& x);
// adjoint.cpp: checkpoint of the state before step s (when enabled)
void adjoint_checkpoint(titus_ctx& x, int s);

struct titus_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  // problem (host copy)
  int nx = 0, ny = 0, C = 0, nm = 0, nq = 0, ne = 0, nsig = 0, integrator = 0, precision = 0;
  double dx = 0, dy = 0, dE = 0, theta = 0, theta_index = 0, rank_tol = 1e-7;
  int rank_cap = 0, capx = 0;
  double cfl = 0.0;           // dE / min(h) * max(1 / rho) (SolverCSD.jl:963)
  std::vector<double> stopping, sigma, bc_scale;
  std::vector<int> degree, cand;
  HMat Ax, Az, aAx, aAz, M;
  // angular operators paired with the spatial stencils, in compressed rows:
  // |Ax|, |Az| (, |Ay|), Ax, Az (, Ay) -- dissipative first, then central
  std::vector<titus::SpMat> sA;
  int nd = 2;                 // spatial dimensions (3: k axis paired with Ay)
  int nz = 1;
  double dz = 1.0;
  int flux_order = 1;         // uncollided flux: 1 upwind, 2 limited (solveFlux!)
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
  double* Kc = nullptr;                     // copy of K before its projection (parallel BUG)
  // lockstep batch (titus_solve_batch_lockstep): the group this context
  // advances with, or null when it runs alone.  In a group the context's
  // launches are recorded in `rec` and issued by the group at the sync
  // points; the pinned mirrors are read and written by copy kernels through
  // their device addresses.
  LockstepGroup* group = nullptr;
  dev::Recorder rec;
  double *h_arena_dev = nullptr, *h_upl_dev = nullptr;
  size_t upl_sent = 0;      // upload ring prefix the group has copied to the device
  double *partial = nullptr, *arena = nullptr, *upl = nullptr;
  int *bcells = nullptr, *cand_d = nullptr;
  unsigned char *bmask = nullptr, *restore = nullptr;
  // pinned host mirrors
  double *h_arena = nullptr, *h_upl = nullptr;
  size_t arena_n = 0, arena_top = 0, upl_n = 0, upl_top = 0;
  size_t fetched_top = 0;                   // arena prefix already copied to h_arena
  hipEvent_t ev_part = nullptr;             // end of the last prefix copy
  dev::Grid g{};
  int nb = 0;
  // state
  HMat S, W;
  int rank = 0, steps_done = 0;
  std::vector<int> ranks;
  std::vector<double> etas;
  mutable Timing timing;
  // running hash of the solve's discrete decisions (kept ranks, completion
  // vectors, truncation ranks): equal hashes of two solves mean the same
  // decisions, i.e. both lie on one smooth branch of the forward map
  uint64_t decision_hash = 1469598103934665603ull;
  int64_t decisions = 0;
  void fold(int64_t v) {
    decision_hash = (decision_hash ^ (uint64_t)v) * 1099511628211ull;
    ++decisions;
  }
  // adjoint: checkpoints and recomputation (adjoint.cpp)
  std::shared_ptr<AdjointStore> adj;
  StepTape* tape = nullptr;
  // angular operators on the device (compressed rows of sA, in its order) and
  // the blocks of the angular products: AW = A_o W (nops x nm x capx), Z of
  // the K-step; the term order of the streaming sums
  dev::AngOps angops{};
  int *ang_rowptr = nullptr, *ang_col = nullptr, *ang_opnz = nullptr, *ang_order = nullptr;
  double *ang_val = nullptr, *AWd = nullptr, *Zd = nullptr;
  // rejection of the rank-adaptive step (SolverCSD.jl:2204-2233, 2282-2301):
  // c_eta > 0 turns it on; the state at the start of each step is kept so
  // that a rejected step restarts from it
  double c_eta = 0.0;
  int64_t rejections = 0;
  double *X_start = nullptr, *dose_start = nullptr, *u0_start = nullptr;
  void* psi_start = nullptr;
  HMat S_start, W_start;
  int r_start = 0;

  ~titus_ctx() {
    for (const CapturedRange& g : range_graphs) (void)hipGraphExecDestroy(g.exec);
    void* ptrs[] = {rho, irho, omega, bc_base, m0, psi, psi2, dose, u0p, X, Xn, K, Kc, Q1, Q2, partial, arena,
                    upl, bcells, cand_d, bmask, restore, X_start, dose_start, u0_start, psi_start,
                    xmid_d, ymid_d, ang_rowptr, ang_col, ang_opnz, ang_order, ang_val, AWd, Zd};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    void* hptrs[] = {h_arena, h_upl};
    for (void* p : hptrs) if (p) (void)hipHostFree(p);
    if (ev_part) (void)hipEventDestroy(ev_part);
    if (st) (void)hipStreamDestroy(st);
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
  void areset() { arena_top = fetched_top = 0; }
  // one transfer of what the arena holds beyond the prefix already copied,
  // then wait for the stream
  void fetch() {
    if (group) {
      // nothing recorded, uploaded or allocated since the last flush: the
      // host already holds every result (a fetch right after wait_part)
      if (!group_pending()) {
        upl_top = upl_sent = 0;
        return;
      }
      lockstep_sync(*this);
      upl_top = upl_sent = 0;
      return;
    }
    if (arena_top > fetched_top)
      HIPCHK(hipMemcpyAsync(h_arena + fetched_top, arena + fetched_top,
                            (arena_top - fetched_top) * sizeof(double), hipMemcpyDeviceToHost, st));
    fetched_top = arena_top;
    HostScope hs(timing, "host:wait");
    HIPCHK(hipStreamSynchronize(st));
    upl_top = 0;          // every queued upload has been consumed
  }
  // in a group: launches, uploads or results waiting for the group's flush
  bool group_pending() const {
    return !rec.ops.empty() || arena_top > fetched_top || upl_top > upl_sent;
  }
  // device-to-device copy on the context's stream (a recorded copy kernel in
  // a group)
  void dcopy(void* dst, const void* src, size_t bytes) {
    if (group) dev::launch_copy(st, dst, src, bytes);
    else HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st));
  }
  // the arena so far goes to the host behind the work queued until now;
  // wait_part() waits for that copy only, so the host can use those results
  // while later work runs on the device
  void fetch_part() {
    // in a group the whole arena reaches the host at the end of every flush
    if (group) return;
    if (arena_top > fetched_top)
      HIPCHK(hipMemcpyAsync(h_arena + fetched_top, arena + fetched_top,
                            (arena_top - fetched_top) * sizeof(double), hipMemcpyDeviceToHost, st));
    fetched_top = arena_top;
    HIPCHK(hipEventRecord(ev_part, st));
  }
  void wait_part() {
    if (group) {
      if (group_pending()) lockstep_sync(*this);
      return;
    }
    HostScope hs(timing, "host:wait");
    HIPCHK(hipEventSynchronize(ev_part));
  }
  HMat view(size_t off, int m, int n) const {
    HMat A(m, n);
    std::memcpy(A.a.data(), h_arena + off, sizeof(double) * (size_t)m * n);
    return A;
  }
  double* upload_raw(const double* src, size_t n) {
    if (group) {
      // the ring is reset only at the group's full syncs: it must hold a phase
      if (upl_top + n > upl_n) throw TitusError(TITUS_E_ARG, "upload ring too small for a batched phase");
      // copied to the device with the group's other uploads before the
      // flush's first kernel
      double* d = upl + upl_top;
      std::memcpy(h_upl + upl_top, src, n * sizeof(double));
      upl_top += (n + 31) & ~size_t(31);
      return d;
    }
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
  // small coefficient blocks travel in the kernel arguments, larger ones
  // through the upload ring
  void gemm(const double* A, int k, const HMat& T, double alpha, double beta, double* out,
            const unsigned char* mask) {
    if (group || T.a.size() > (size_t)dev::kArgMat) {
      gemm_dev(A, k, upload(T), T.n, alpha, beta, out, mask);
      return;
    }
    timed("gemm", 8.0 * C * (k + T.n * (beta != 0.0 ? 2 : 1)), 2.0 * C * k * T.n, [&] {
      dev::launch_gemm_args(st, C, A, dev::DT::F64, C, k, T.a.data(), T.n, alpha, beta, out, C,
                            mask, false);
    });
  }

  // ---- host building blocks (angular side) -------------------------------------
  // [A_o Wm for every angular operator]: computed once per basis and reused
  // by the Galerkin products, the L-step (A (W S^T) = (A W) S^T) and the
  // augmented basis (A [W Wt] = [A W, A Wt])
  using AWs = std::vector<HMat>;
  // f(o) for every angular operator o: on one thread per operator when the
  // products are large enough to pay for the threads (the 3D operators) and
  // the context is not one of a batch (whose contexts already run side by
  // side); each product is computed the same way either way
  bool host_parallel = true;
  template <class F>
  void for_ops(size_t work, F&& f) const {
    const int n = (int)sA.size();
    if (!host_parallel || work < (size_t(1) << 20) || n < 2) {
      for (int o = 0; o < n; ++o) f(o);
      return;
    }
    std::vector<std::exception_ptr> err(n);
    std::vector<std::thread> th;
    for (int o = 1; o < n; ++o)
      th.emplace_back([&f, &err, o] {
        try { f(o); } catch (...) { err[o] = std::current_exception(); }
      });
    try { f(0); } catch (...) { err[0] = std::current_exception(); }
    for (auto& t : th) t.join();
    for (auto& e : err)
      if (e) std::rethrow_exception(e);
  }
  AWs apply_a(const HMat& Wm) const {
    HostScope hs(timing, "host:galerkin_w");
    AWs out(sA.size());
    for_ops(sA.empty() ? 0 : sA[0].val.size() * (size_t)Wm.n,
            [&](int o) { out[o] = titus::spmm(sA[o], Wm); });
    return out;
  }
  std::vector<HMat> galerkin_from(const HMat& Wa, const AWs& awb) const {
    HostScope hs(timing, "host:galerkin_w");
    std::vector<HMat> out(awb.size());
    for_ops((size_t)Wa.m * Wa.n * (awb.empty() ? 0 : awb[0].n),
            [&](int o) { out[o] = titus::matmul(Wa, awb[o], true); });
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
  HMat range_w(const HMat& L, int ncols, const HMat* against, double scale,
               titus::OrthRecord* rec = nullptr) {
    HostScope hs(timing, "host:range_w");
    static const std::vector<int> all;
    int p = 0;
    std::vector<int> chosen;
    HMat B = titus::range_basis(L, ncols, against, scale, rank_tol, true, all, &p, rec, &chosen);
    fold(p);
    for (int c : chosen) fold(c);
    return B;
  }
  void stream_k(int r, const std::vector<HMat>& WAW) {
    const HMat Z = stream_Z(S, WAW);
    if (!group && Z.a.size() <= (size_t)dev::kArgMat) {
      timed("stream_K", 8.0 * C * (2 * r + 1), 2.0 * (2 * nd + 2) * C * r * r,
            [&] { dev::launch_stream_k_args(st, g, X, C, r, Z.a.data(), r, dE, K, C, full32); });
      return;
    }
    const double* Zd = upload(Z);
    timed("stream_K", 8.0 * C * (2 * r + 1), 2.0 * (2 * nd + 2) * C * r * r,
          [&] { dev::launch_stream_k(st, g, X, C, r, Zd, r, dE, K, C, full32); });
  }

  // ---- angular products on the device ---------------------------------------------
  // The operators' compressed rows and the term order, copied once (titus_create).
  void setup_angular() {
    const int nops_ = (int)sA.size();
    std::vector<int> rp, cl, nz;
    std::vector<double> vl;
    for (const titus::SpMat& A : sA) {
      nz.push_back((int)cl.size());
      rp.insert(rp.end(), A.rowptr.begin(), A.rowptr.end());
      cl.insert(cl.end(), A.col.begin(), A.col.end());
      vl.insert(vl.end(), A.val.begin(), A.val.end());
    }
    const std::vector<int> ord = term_order();
    ang_rowptr = dalloc<int>(rp.size());
    ang_col = dalloc<int>(cl.size());
    ang_val = dalloc<double>(vl.size());
    ang_opnz = dalloc<int>(nz.size());
    ang_order = dalloc<int>(ord.size());
    HIPCHK(hipMemcpy(ang_rowptr, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice));
    if (!cl.empty()) {
      HIPCHK(hipMemcpy(ang_col, cl.data(), sizeof(int) * cl.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(ang_val, vl.data(), sizeof(double) * vl.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(ang_opnz, nz.data(), sizeof(int) * nz.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ang_order, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice));
    angops = dev::AngOps{nm, nops_, ang_rowptr, ang_col, ang_val, ang_opnz};
    AWd = dalloc<double>((size_t)nops_ * nm * capx);
    Zd = dalloc<double>((size_t)capx * (2 * nd + 2) * capx);
  }
  // arena offset of [W^T A_o W for every operator] for a device copy Wdev of W
  // (nm x k); AWd receives A_o W
  size_t galerkin_w_dev(const double* Wdev, int k) {
    const size_t off = aalloc((size_t)nops() * k * k);
    timed("angular", 8.0 * nm * k * (1 + nops()), 2.0 * nops() * k * (angops_nnz() + (double)nm * k),
          [&] { dev::launch_ang_apply(st, angops, Wdev, k, Wdev, k, AWd, arena + off); });
    return off;
  }
  double angops_nnz() const {
    double n = 0.0;
    for (const titus::SpMat& A : sA) n += (double)A.val.size();
    return n / std::max<size_t>(sA.size(), 1);
  }
  // The streaming K-step with its coefficients formed on the device from the
  // Galerkin blocks at arena offset gw (W^T A_o W of the current W)
  void stream_k_dev(int r, const double* Sdev, size_t gw) {
    timed("angular", 8.0 * (2.0 * nops() * r * r + (2 * nd + 2) * r * r), 2.0 * (2 * nd + 1) * r * r * r,
          [&] { dev::launch_ang_stream_z(st, r, r, nd, dx, dy, dz, Sdev, arena + gw, Zd); });
    timed("stream_K", 8.0 * C * (2 * r + 1), 2.0 * (2 * nd + 2) * C * r * r,
          [&] { dev::launch_stream_k(st, g, X, C, r, Zd, r, dE, K, C, full32); });
  }
  // arena offset of the L-step L = W S^T - dE sum_o A_o W S^T (X^T L_o X)^T
  // (nm x r) from AWd (of this W) and the Galerkin blocks at arena offset gx
  size_t stream_l_dev(int r, const double* Wdev, const double* Sdev, size_t gx) {
    const size_t off = aalloc((size_t)nm * r);
    timed("angular", 8.0 * nm * r * (2 + nops()), 2.0 * nops() * nm * r * r,
          [&] { dev::launch_ang_stream_l(st, nm, r, nops(), ang_order, dE, Wdev, Sdev, AWd,
                                         arena + gx, arena + off); });
    return off;
  }

  // ---- device range basis --------------------------------------------------------
  // The device part of a range basis is the same ~20 launches whenever its
  // (K block, basis block, widths, arena position) recur, which is every step
  // at a steady rank: the sequence is captured once into a graph and then
  // replayed by one launch.  Every argument of those kernels is a device
  // address or a constant (the coefficient blocks are arena results), so a
  // replay computes what the launches compute.  Not in a lockstep group (its
  // launches are recorded), on timed steps (events around the launches) or
  // while the adjoint tapes a step.
  struct CapturedRange {
    double* Kw;
    const double* Xa;
    int k, q;
    size_t top0, top1;     // arena position before and after the job
    RangeJob job;
    hipGraphExec_t exec;
  };
  std::vector<CapturedRange> range_graphs;   // oldest first
  static constexpr size_t kMaxRangeGraphs = 8;

  RangeJob range_launch(double* Kw, int k, const double* Xa, int q) {
    if (group || dev::t_recorder || timing.enabled || tape) return range_launch_direct(Kw, k, Xa, q);
    const size_t top0 = arena_top;
    for (const CapturedRange& g : range_graphs)
      if (g.Kw == Kw && g.Xa == Xa && g.k == k && g.q == q && g.top0 == top0) {
        HIPCHK(hipGraphLaunch(g.exec, st));
        arena_top = g.top1;
        return g.job;
      }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    RangeJob j;
    try {
      j = range_launch_direct(Kw, k, Xa, q);
    } catch (...) {
      (void)hipStreamEndCapture(st, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    HIPCHK(hipStreamEndCapture(st, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    if (range_graphs.size() >= kMaxRangeGraphs) {
      (void)hipGraphExecDestroy(range_graphs.front().exec);
      range_graphs.erase(range_graphs.begin());
    }
    range_graphs.push_back(CapturedRange{Kw, Xa, k, q, top0, arena_top, j, exec});
    HIPCHK(hipGraphLaunch(exec, st));
    return j;
  }

  // Launch the device part of the basis of range(Kw) minus span(Xa): the
  // projection (twice), three Gram + shifted-Cholesky passes and the candidate
  // rows; everything the host needs lands in the arena.  Kw is overwritten.
  RangeJob range_launch_direct(double* Kw, int k, const double* Xa, int q) {
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
  int range_finish(const RangeJob& j, int ncols, double* out, SpatialOrthTape* ot = nullptr) {
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
    if (ot) ot->k = k;
    if (tr1 > 0.0) {
      for (int t = 0; t < 3; ++t)
        if (h_arena[j.st + t] < 0.0) throw TitusError(TITUS_E_RANK, "shifted Cholesky failed");
      HMat R1 = view(j.r1, k, k), R2 = view(j.r2, k, k), R3 = view(j.r3, k, k);
      HMat R = titus::matmul(R3, titus::matmul(R2, R1));
      HMat P, Zr;
      std::vector<double> sv;
      titus::svd(R, P, sv, Zr, ot != nullptr);
      const double sc = scale < 0.0 ? (sv.empty() ? 0.0 : sv[0]) : scale;
      if (sc > 0.0)
        for (double x : sv) if (x > rank_tol * sc) ++p;
      p = std::min(p, ncols);
      HMat R3inv = titus::inv_upper(R3);
      T = titus::matmul(R3inv, titus::cols(P, 0, p));                    // k x p
      if (ot) {
        ot->sv = sv; ot->P = P; ot->V = Zr; ot->R3inv = R3inv;
        dcopy(ot->Q2, Q2, sizeof(double) * (size_t)C * k);
      }
      HMat Uc = titus::matmul(view(j.rows_q, j.nc, k), T);               // nc x p
      Bc = j.Xa ? titus::hcat(view(j.rows_x, j.nc, q), Uc) : Uc;
    } else {
      Bc = j.Xa ? view(j.rows_x, j.nc, q) : HMat(j.nc, 0);
    }
    const int need = ncols - p;
    if (ot) {
      ot->p = p;
      ot->Tp = HMat(p, ncols);
      for (int i = 0; i < p; ++i) ot->Tp(i, i) = 1.0;
    }
    std::vector<int> chosen;
    HMat Rinv;
    if (need > 0 && !titus::canonical_completion(Bc, cidx, need, false, chosen, Rinv))
      throw TitusError(TITUS_E_RANK, "canonical completion ran out of candidates");
    fold(p);
    for (int c : chosen) fold(c);
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
      HMat UR = titus::matmul(Uch, Rinv, true, false);                    // p x need
      if (ot)
        for (int s = 0; s < need; ++s)
          for (int c = 0; c < p; ++c) ot->Tp(c, p + s) = -UR(c, s);
      HMat Tc = titus::matmul(T, UR);                                      // k x need
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
    if (tape)
      dcopy(tape->psi_a, psi, sizeof(double) * (size_t)C * nq);
    dose_half(stopping[s - 1]);
    const double denom = 1.0 + dE * sg[0];
    // psi read and written once, density, restore flags and moment partials
    timed("sweep", 2.0 * ps * C * nq + 9.0 * C + 8.0 * C * u0_groups, 14.0 * C * nq, [&] {
      dev::launch_sweep(st, g, psi, psi2, tpsi, nq, omega, dE, denom, restore, m0, u0p, sweep_plan);
    });
    std::swap(psi, psi2);
    if (tape)
      dcopy(tape->psi_b, psi, sizeof(double) * (size_t)C * nq);
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
    if (tape) { tape->Ws = W; tape->Ss = S; tape->Lo = L; }
    W = range_w(L, rank, nullptr, -1.0, tape ? &tape->Rw2 : nullptr);
    S = titus::matmul(L, W, true, false);
    if (tape) { tape->W2 = W; tape->S3 = S; }
  }
  HMat mtdw(const std::vector<double>& Dvec, const HMat& Wm) const {
    HMat DW = Wm;
    for (int j = 0; j < DW.n; ++j)
      for (int i = 0; i < nm; ++i) DW(i, j) *= Dvec[i];
    return titus::matmul(M, DW, true, false);                  // nq x r
  }
  // K = X S + dE psi MtDW with ghost rows zeroed (SolverCSD.jl:1079-1083)
  // (mask: the ghost rows zeroed, as SolverCSD.jl:1083; NULL keeps them, as
  // the CUDA.jl variant does, :1651)
  void scatter_K(const HMat& MtDW, const unsigned char* mask) {
    gemm(X, rank, S, 1.0, 0.0, K, nullptr);
    HMat T = MtDW;
    for (double& v : T.a) v *= dE;
    const double ps = mixed ? 4.0 : 8.0;
    if (!group && T.a.size() <= (size_t)dev::kArgMat) {
      timed("gemm", ps * C * nq + 16.0 * C * T.n, 2.0 * C * nq * T.n, [&] {
        dev::launch_gemm_args(st, C, psi, tpsi, C, nq, T.a.data(), T.n, 1.0, 1.0, K, C, mask,
                              full32);
      });
      return;
    }
    const double* Td = upload(T);
    timed("gemm", ps * C * nq + 16.0 * C * T.n, 2.0 * C * nq * T.n, [&] {
      dev::launch_gemm(st, C, psi, tpsi, C, nq, Td, T.n, 1.0, 1.0, K, C, mask, full32);
    });
  }
  void scatter_K(const HMat& MtDW) { scatter_K(MtDW, bmask); }
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

  // One energy step of the fixed-rank integrator.  With `tape` set, every
  // intermediate the adjoint needs is recorded on the way (copies only; the
  // arithmetic is the same).
  void step_fixed(int s) {
    std::vector<double> Dvec, Dout;
    const size_t cr = sizeof(double) * (size_t)C * rank;
    if (tape) {
      tape->s = s; tape->S = S; tape->W = W;
      dcopy(tape->X, X, cr);
    }
    begin_step(s, Dvec, Dout);
    if (tape) { tape->Dvec = Dvec; tape->Dout = Dout; tape->sig0 = sig(s)[0]; }
    const int r = rank;
    size_t xtp0 = 0;
    if (s > 1) {
      // K-step and old Galerkin products (SolverCSD.jl:1019-1038)
      // the Galerkin products of X and W, the L-step (:1033-1046) and the
      // K-step coefficients on the device; the host takes L as soon as it is
      // there while the device runs the K-step and its range job
      areset();
      const size_t gx = galerkin_x(X, r, X, r);
      const double* Wdev = upload(W);
      const double* Sdev = upload(S);
      const size_t gw = galerkin_w_dev(Wdev, r);
      const size_t lo = stream_l_dev(r, Wdev, Sdev, gx);
      fetch_part();
      stream_k_dev(r, Sdev, gw);
      RangeJob j = range_launch(K, r, nullptr, 0);
      wait_part();
      const std::vector<HMat> WAW = galerkin_view(gw, r, r);
      std::vector<HMat> XLX = galerkin_view(gx, r, r);
      HMat L = view(lo, nm, r);
      HMat WNew = range_w(L, r, nullptr, -1.0, tape ? &tape->Rw1 : nullptr);
      HMat NUp = titus::matmul(WNew, W, true, false);
      fetch();
      range_finish(j, r, Xn, tape ? &tape->Rx1 : nullptr);
      // the device writes XNew and the Galerkin products of XNew and WNew
      areset();
      const size_t mu = tn(Xn, r, X, r, dev::OP_NONE);
      const size_t gn = galerkin_x(Xn, r, Xn, r);
      const size_t gw1 = galerkin_w_dev(upload(WNew), r);
      fetch();
      const std::vector<HMat> WAW1 = galerkin_view(gw1, r, r);
      HMat MUp = view(mu, r, r);
      if (tape) {
        tape->WAW = WAW; tape->XLX = XLX; tape->L0 = titus::matmul(W, S, false, true);
        tape->W1 = WNew; tape->NUp1 = NUp; tape->MUp1 = MUp;
      }
      XLX = galerkin_view(gn, r, r);
      W = WNew;
      std::swap(X, Xn);
      // X^T psi of the in-scattering needs only the new X: it runs on the
      // device during the host S-step and out-scattering
      areset();
      xtp0 = tn_psi(X, r);
      fetch_part();
      // S-step (:1051-1063)
      S = titus::matmul(titus::matmul(MUp, S), NUp, false, true);
      if (tape) {
        tape->S1 = S; tape->XLX1 = XLX; tape->WAW1 = WAW1;
        dcopy(tape->X1, X, cr);
      }
      S = stream_S(S, XLX, WAW1);
    } else {
      areset();
      xtp0 = tn_psi(X, r);
      fetch_part();
    }
    out_scatter(Dout);
    in_scatter_fixed(Dvec, xtp0, bmask);
    dose_half(stopping[s]);
  }

  // In-scattering of the fixed-rank step (SolverCSD.jl:1079-1102; the CUDA.jl
  // variant :1649-1676 with mask = NULL): K = X S + dE psi M^T D W, new X from
  // range(K), L = W S^T + dE D M (psi^T X), new W from range(L), then
  // S = MUp S NUp^T + dE (X^T psi) M^T D W.  xtp0: arena offset of X^T psi
  // (old X), fetched in part before the call.
  void in_scatter_fixed(const std::vector<double>& Dvec, size_t xtp0, const unsigned char* mask) {
    const int r = rank;
    const size_t cr = sizeof(double) * (size_t)C * r;
    const HMat MtDW = mtdw(Dvec, W);
    scatter_K(MtDW, mask);
    RangeJob j = range_launch(K, r, nullptr, 0);
    // the new W on the host while the device runs the K range job
    wait_part();
    HMat PtX = titus::transpose(view(xtp0, r, nq));
    HMat L = titus::matmul(W, S, false, true);
    HMat MP = titus::matmul(M, PtX);
    for (int c = 0; c < r; ++c)
      for (int i = 0; i < nm; ++i) L(i, c) += dE * Dvec[i] * MP(i, c);
    HMat WNew = range_w(L, r, nullptr, -1.0, tape ? &tape->Rw3 : nullptr);
    HMat NUp = titus::matmul(WNew, W, true, false);
    fetch();
    range_finish(j, r, Xn, tape ? &tape->Rx3 : nullptr);
    // products of XNew on the device while the host forms M^T D WNew
    areset();
    const size_t mu = tn(Xn, r, X, r, dev::OP_NONE);
    const size_t xtp = tn_psi(Xn, r);
    const HMat B3 = mtdw(Dvec, WNew);
    fetch();
    HMat MUp = view(mu, r, r);
    HMat XtP = view(xtp, r, nq);
    W = WNew;
    std::swap(X, Xn);
    S = titus::matmul(titus::matmul(MUp, S), NUp, false, true);
    HMat add = titus::matmul(XtP, B3);
    for (size_t i = 0; i < S.a.size(); ++i) S.a[i] += dE * add.a[i];
    if (tape) {
      tape->MtDW = MtDW; tape->MUp3 = MUp; tape->W3 = W; tape->NUp3 = NUp; tape->B3 = B3;
      tape->A3 = XtP; tape->S5 = S;
      dcopy(tape->X3, X, cr);
    }
  }

  // ---- the CUDA.jl variant: CudaSolveFirstCollisionSourceDLR4thOrder ----------
  // (SolverCSD.jl:1422-1724): fixed rank; the streaming K-, L- and S-steps
  // integrate the central terms only (L2x with Ax, L2y with Az) with RK4; the
  // uncollided flux is the analytic beam (k_analytic_psi) instead of a sweep.
  bool beam_set = false;
  double beam_x0 = 0, beam_y0 = 0, beam_om1 = 0, beam_om3 = 0, int_sigma = 0;
  double *xmid_d = nullptr, *ymid_d = nullptr;
  std::vector<double> tau;
  // classical RK4 of Y' = f(Y) over dE on the host (the reference's stages)
  template <class F>
  HMat rk4_host(const HMat& Y, F&& f) const {
    const double h = 0.5 * dE;
    auto axpy = [](const HMat& a, double c, const HMat& b) {
      HMat o = a;
      for (size_t i = 0; i < o.a.size(); ++i) o.a[i] += c * b.a[i];
      return o;
    };
    HMat k = f(Y);
    HMat up = axpy(Y, dE / 6, k);
    k = f(axpy(Y, h, k));
    up = axpy(up, 2 * dE / 6, k);
    k = f(axpy(Y, h, k));
    up = axpy(up, 2 * dE / 6, k);
    k = f(axpy(Y, dE, k));
    return axpy(up, dE / 6, k);
  }
  // K <- RK4 of K' = -L2x K Bx - L2y K Bz on the device: each f is one
  // stencil launch (centre and dissipative blocks zero, dE = 1), the stage
  // sums are copies and axpys on C x r blocks (Q1: the running sum, Q2: the
  // stage argument, Kc: the stage value)
  void rk4_k_device(int r, const HMat& Bx, const HMat& Bz) {
    HMat Z(r, 6 * r);
    for (int b = 0; b < r; ++b)
      for (int a = 0; a < r; ++a) {
        Z(a, 2 * r + b) = Bx(a, b) / (2.0 * dx);
        Z(a, 3 * r + b) = -Bx(a, b) / (2.0 * dx);
        Z(a, 4 * r + b) = Bz(a, b) / (2.0 * dy);
        Z(a, 5 * r + b) = -Bz(a, b) / (2.0 * dy);
      }
    const double* Zd = Z.a.size() <= (size_t)dev::kArgMat ? nullptr : upload(Z);
    const size_t n = (size_t)C * r, bytes = sizeof(double) * n;
    auto f = [&](const double* Y) {
      timed("stream_K", 8.0 * C * (2 * r + 1), 8.0 * C * r * r, [&] {
        if (Zd) dev::launch_stream_k(st, g, Y, C, r, Zd, r, 1.0, Kc, C, false);
        else dev::launch_stream_k_args(st, g, Y, C, r, Z.a.data(), r, 1.0, Kc, C, false);
      });
    };
    auto copy = [&](double* dst, const double* src) {
      dcopy(dst, src, bytes);
    };
    const double h = 0.5 * dE;
    f(K);
    copy(Q1, K); dev::launch_axpy(st, n, dE / 6, Kc, Q1);
    copy(Q2, K); dev::launch_axpy(st, n, h, Kc, Q2);
    f(Q2);
    dev::launch_axpy(st, n, 2 * dE / 6, Kc, Q1);
    copy(Q2, K); dev::launch_axpy(st, n, h, Kc, Q2);
    f(Q2);
    dev::launch_axpy(st, n, 2 * dE / 6, Kc, Q1);
    copy(Q2, K); dev::launch_axpy(st, n, dE, Kc, Q2);
    f(Q2);
    copy(K, Q1); dev::launch_axpy(st, n, dE / 6, Kc, K);
  }
  void step_rk4(int s) {
    if (!beam_set) throw TitusError(TITUS_E_STATE, "RK4 variant: set the analytic beam before the first step");
    const std::vector<double> sg = sig(s);
    dose_half(stopping[s - 1]);
    int_sigma += dE * sg[0];
    const double ps = mixed ? 4.0 : 8.0;
    timed("psi", ps * C * nq + 8.0 * C, 20.0 * C * nq, [&] {
      dev::launch_analytic_psi(st, g, psi, tpsi, nq, omega, xmid_d, ymid_d, beam_x0, beam_y0,
                               beam_om1, beam_om3, tau[s], int_sigma, m0, u0p);
    });
    std::vector<double> Dvec(nm), Dout(nm);
    for (int i = 0; i < nm; ++i) {
      Dvec[i] = sg[degree[i]];
      Dout[i] = sg[0] - Dvec[i];
    }
    const int r = rank;
    size_t xtp0 = 0;
    if (s > 1) {
      // K-step (:1580-1597) with X^T L X of the old X on the device meanwhile
      areset();
      const size_t gx = galerkin_x(X, r, X, r);
      fetch_part();
      const AWs aw = apply_a(W);                      // [|Ax| W, |Az| W, Ax W, Az W]
      gemm(X, r, S, 1.0, 0.0, K, nullptr);
      rk4_k_device(r, titus::matmul(W, aw[2], true, false), titus::matmul(W, aw[3], true, false));
      RangeJob j = range_launch(K, r, nullptr, 0);
      // L-step (:1600-1617) on the host with the old X
      wait_part();
      std::vector<HMat> XLX = galerkin_view(gx, r, r);
      const HMat& Lx = XLX[2];
      const HMat& Lz = XLX[3];
      HMat L = rk4_host(titus::matmul(W, S, false, true), [&](const HMat& Y) {
        HMat a = titus::matmul(titus::spmm(sA[2], Y), Lx, false, true);
        HMat b = titus::matmul(titus::spmm(sA[3], Y), Lz, false, true);
        for (size_t i = 0; i < a.a.size(); ++i) a.a[i] = -a.a[i] - b.a[i];
        return a;
      });
      HMat WNew = range_w(L, r, nullptr, -1.0);
      HMat NUp = titus::matmul(WNew, W, true, false);
      fetch();
      range_finish(j, r, Xn);
      areset();
      const size_t mu = tn(Xn, r, X, r, dev::OP_NONE);
      const size_t gn = galerkin_x(Xn, r, Xn, r);
      const AWs awn = apply_a(WNew);
      fetch();
      HMat MUp = view(mu, r, r);
      XLX = galerkin_view(gn, r, r);
      W = WNew;
      std::swap(X, Xn);
      // S-step (:1621-1634)
      S = titus::matmul(titus::matmul(MUp, S), NUp, false, true);
      const HMat Bx = titus::matmul(W, awn[2], true, false), Bz = titus::matmul(W, awn[3], true, false);
      S = rk4_host(S, [&](const HMat& Y) {
        HMat a = titus::matmul(titus::matmul(XLX[2], Y), Bx);
        HMat b = titus::matmul(titus::matmul(XLX[3], Y), Bz);
        for (size_t i = 0; i < a.a.size(); ++i) a.a[i] = -a.a[i] - b.a[i];
        return a;
      });
    }
    areset();
    xtp0 = tn_psi(X, r);
    fetch_part();
    out_scatter(Dout);
    in_scatter_fixed(Dvec, xtp0, nullptr);
    dose_half(stopping[s]);
  }

  // truncate! (SolverCSD.jl:3369-3399) of the 2r x 2r Sup: the SVD and the
  // kept rank (plan_truncation), then the new factors over the augmented
  // bases X1 (device, C x 2r) and W1 (apply_truncation: X (device), W, S,
  // rank).  Split so that a rejected step can look at the decision first.
  struct TruncPlan {
    HMat U, V;
    std::vector<double> D;
    int rmax = 0;
    double kept_norm = 0.0;   // |S| of the truncated S (Frobenius)
  };
  TruncPlan plan_truncation(const HMat& Sup) {
    HostScope hs(timing, "host:truncate");
    TruncPlan t;
    titus::svd_completed(Sup, rank_tol, t.U, t.D, t.V);
    const std::vector<double>& D = t.D;
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
    t.rmax = std::max(std::min(rmax, rank_cap), 2);
    for (int i = 0; i < t.rmax; ++i) t.kept_norm += D[i] * D[i];
    t.kept_norm = std::sqrt(t.kept_norm);
    return t;
  }
  void apply_truncation(const TruncPlan& t, const double* X1, const HMat& W1) {
    HostScope hs(timing, "host:truncate");
    const int n = (int)t.D.size(), rmax = t.rmax;
    fold(rmax);
    gemm(X1, n, titus::cols(t.U, 0, rmax), 1.0, 0.0, X, nullptr);
    W = titus::matmul(W1, titus::cols(t.V, 0, rmax));
    S = HMat(rmax, rmax);
    for (int i = 0; i < rmax; ++i) S(i, i) = t.D[i];
    rank = rmax;
  }
  void truncate_to(const HMat& Sup, const double* X1, const HMat& W1) {
    apply_truncation(plan_truncation(Sup), X1, W1);
  }

  // ---- rejection (SolverCSD.jl:2204-2233, 2282-2301) ---------------------------
  // State at the start of the step (device copies of X, psi, dose and the
  // moment partials; host S, W, rank).
  void save_start() {
    const size_t psz = mixed ? sizeof(float) : sizeof(double);
    dcopy(X_start, X, sizeof(double) * (size_t)C * rank);
    dcopy(psi_start, psi, psz * (size_t)C * nq);
    dcopy(dose_start, dose, sizeof(double) * C);
    dcopy(u0_start, u0p, sizeof(double) * (size_t)u0_groups * C);
    S_start = S;
    W_start = W;
    r_start = rank;
  }
  // psi, dose and the moment partials back to the step's start; S padded
  // with zeros to the new rank (u = X S W^T is unchanged)
  void restart_state(int newrank, size_t neta) {
    const size_t psz = mixed ? sizeof(float) : sizeof(double);
    dcopy(psi, psi_start, psz * (size_t)C * nq);
    dcopy(dose, dose_start, sizeof(double) * C);
    dcopy(u0p, u0_start, sizeof(double) * (size_t)u0_groups * C);
    S = HMat(newrank, newrank);
    for (int j = 0; j < r_start; ++j)
      for (int i = 0; i < r_start; ++i) S(i, j) = S_start(i, j);
    rank = newrank;
    ranks.pop_back();
    etas.resize(neta);
    ++rejections;
    fold(-1 - newrank);
  }
  // streaming half rejected: X1 = [X_start | Xt] (device Xn), W1 = [W_start | Wt]
  void reject_streaming(const HMat& W1, size_t neta) {
    const int n = W1.n;
    dcopy(X, Xn, sizeof(double) * (size_t)C * n);
    W = W1;
    restart_state(n, neta);
  }
  // scattering half rejected: its rc new directions (Xn columns rc..2rc-1,
  // W1 columns rc..) made orthogonal to the step's starting bases
  void reject_scattering(int rc, const HMat& W1, size_t neta) {
    const int r0 = r_start;
    dcopy(K, Xn + (size_t)rc * C, sizeof(double) * (size_t)C * rc);
    areset();
    RangeJob j = range_launch(K, rc, X_start, r0);
    const HMat Wt = titus::cols(W1, rc, W1.n);
    HMat Wt2 = range_w(Wt, rc, &W_start, titus::frob(Wt));
    fetch();
    dcopy(X, X_start, sizeof(double) * (size_t)C * r0);
    range_finish(j, rc, X + (size_t)r0 * C);
    W = titus::hcat(W_start, Wt2);
    restart_state(r0 + rc, neta);
  }
  HMat augment_w(const HMat& L) {
    return titus::hcat(W, range_w(L, rank, &W, titus::frob(L)));
  }
  void copy_x_to_aug(int r) {
    dcopy(Xn, X, (size_t)C * r * sizeof(double));
  }

  void step_bug(int s) {
    // a rejected attempt restores the step's start and enlarges the rank;
    // every rejection needs 2r < cap and adds r >= 1 columns, so the loop
    // ends
    for (;;) {
      if (c_eta > 0.0) save_start();
      if (step_bug_attempt(s)) return;
    }
  }

  // One attempt at a BUG step; false when the rejection test sent it back
  // to its start (state already restored and enlarged).
  bool step_bug_attempt(int s) {
    ranks.push_back(rank);
    const size_t neta = etas.size();
    std::vector<double> Dvec, Dout;
    begin_step(s, Dvec, Dout);
    if (s > 1) {
      const int r = rank;
      // Galerkin products of X and W, the L-step (old X) and the K-step
      // coefficients on the device; the new angular basis on the host as soon
      // as L is there, while the device runs the K-step and its range job
      areset();
      const size_t gx = galerkin_x(X, r, X, r);
      const double* Wdev = upload(W);
      const double* Sdev = upload(S);
      const size_t gw = galerkin_w_dev(Wdev, r);
      const size_t lo = stream_l_dev(r, Wdev, Sdev, gx);
      fetch_part();
      stream_k_dev(r, Sdev, gw);
      copy_x_to_aug(r);
      RangeJob j = range_launch(K, r, X, r);
      wait_part();
      HMat W1 = augment_w(view(lo, nm, r));
      HMat Nu = titus::matmul(W1, W, true, false);
      fetch();
      range_finish(j, r, Xn + (size_t)r * C);               // Xn = X1 = [X | Xt]
      // products of X1 and W1^T A W1 on the device
      areset();
      const size_t mu = tn(Xn, 2 * r, X, r, dev::OP_NONE);
      const size_t g1 = galerkin_x(Xn, 2 * r, Xn, 2 * r);
      const size_t gw1 = galerkin_w_dev(upload(W1), 2 * r);
      fetch();
      std::vector<HMat> WAW = galerkin_view(gw1, 2 * r, 2 * r);
      HMat Mu = view(mu, 2 * r, r);
      std::vector<HMat> XLX = galerkin_view(g1, 2 * r, 2 * r);
      HMat S1 = titus::matmul(titus::matmul(Mu, S), Nu, false, true);   // 2r x 2r
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
      const double eta = titus::frob(acc);
      etas.push_back(eta);
      const TruncPlan tp = plan_truncation(Sup);
      if (c_eta > 0.0 && 2 * r < rank_cap) {
        // rejection test of the streaming half (:2204-2233)
        const double bound = c_eta * theta * std::max(1e-7, std::pow(titus::frob(Sup), theta_index)) / dE;
        if (tp.rmax == 2 * r || eta > bound) {
          reject_streaming(W1, neta);
          return false;
        }
      }
      apply_truncation(tp, Xn, W1);
    }
    // X^T psi (old X) on the device during the host out-scattering
    const int r = rank;
    areset();
    const size_t xtp = tn_psi(X, r);
    fetch_part();
    out_scatter(Dout);
    // in-scattering (:2252-2306)
    scatter_K(mtdw(Dvec, W));
    copy_x_to_aug(r);
    RangeJob j = range_launch(K, r, X, r);
    // the L-step and the new angular basis on the host while the device runs
    // the scattering and the K range job
    wait_part();
    HMat Xtpsi = view(xtp, r, nq);
    HMat L = titus::matmul(W, S, false, true);
    HMat MX = titus::matmul(M, Xtpsi, false, true);
    for (int c = 0; c < r; ++c)
      for (int i = 0; i < nm; ++i) L(i, c) += dE * Dvec[i] * MX(i, c);
    HMat W1 = augment_w(L);
    HMat Nu = titus::matmul(W1, W, true, false);
    fetch();
    range_finish(j, r, Xn + (size_t)r * C);
    // products of X1 on the device while the host forms M^T D W1
    areset();
    const size_t mu = tn(Xn, 2 * r, X, r, dev::OP_NONE);
    const size_t x1p = tn_psi(Xn, 2 * r);
    const HMat MtDW1 = mtdw(Dvec, W1);
    fetch();
    HMat Mu = view(mu, 2 * r, r);
    HMat S1 = titus::matmul(titus::matmul(Mu, S), Nu, false, true);
    HMat X1psi = view(x1p, 2 * r, nq);
    HMat add = titus::matmul(X1psi, MtDW1);
    for (size_t i = 0; i < S1.a.size(); ++i) S1.a[i] += dE * add.a[i];
    const TruncPlan tp = plan_truncation(S1);
    if (c_eta > 0.0 && 2 * r < rank_cap && r_start + r <= rank_cap) {
      // rejection test of the scattering half (:2282-2301): eta =
      // |(Xt^T psi)(M^T D Wt)| from rows r.. of X1^T psi, columns r.. of M^T D W1
      HMat xt(r, nq), mt(nq, r);
      for (int c = 0; c < nq; ++c)
        for (int i = 0; i < r; ++i) xt(i, c) = X1psi(r + i, c);
      for (int c = 0; c < r; ++c)
        for (int i = 0; i < nq; ++i) mt(i, c) = MtDW1(i, r + c);
      const double eta_s = titus::frob(titus::matmul(xt, mt));
      const double bound = c_eta * theta * std::max(1e-7, std::pow(tp.kept_norm, theta_index)) / dE;
      if (tp.rmax == 2 * r || eta_s > bound) {
        reject_scattering(r, W1, neta);
        return false;
      }
    }
    apply_truncation(tp, Xn, W1);
    dose_half(stopping[s]);
    return true;
  }

  // [[S, L^T Wt], [Xt^T K, 0]] (SolverCSD.jl:1882-1888, 1935-1940)
  static HMat augmented_s(const HMat& S0, const HMat& XtK, const HMat& L, const HMat& W1) {
    const int r = S0.m;
    HMat Sn(2 * r, 2 * r);
    const HMat LW = titus::matmul(L, titus::cols(W1, r, W1.n), true, false);   // L^T Wt
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < r; ++i) {
        Sn(i, j) = S0(i, j);
        Sn(r + i, j) = XtK(i, j);
        Sn(i, r + j) = LW(i, j);
      }
    return Sn;
  }

  // Parallel BUG (SolveFirstCollisionSourceDLRParallel, SolverCSD.jl:1726-2008):
  // K-, L- and S-steps all start from the old factors and the old Galerkin
  // products; the new directions Xt, Wt augment the bases and the coupling
  // blocks Xt^T K, L^T Wt fill the augmented coefficient matrix, which
  // truncate! reduces.
  void step_parallel(int s) {
    std::vector<double> Dvec, Dout;
    begin_step(s, Dvec, Dout);
    if (s > 1) {
      const int r = rank;
      areset();
      const size_t gx = galerkin_x(X, r, X, r);
      fetch_part();
      const AWs aw = apply_a(W);
      const std::vector<HMat> WAW = galerkin_from(W, aw);
      stream_k(r, WAW);                                   // K (:1843-1846)
      copy_x_to_aug(r);
      dcopy(Kc, K, sizeof(double) * (size_t)C * r);
      RangeJob j = range_launch(K, r, X, r);              // K is left projected off X
      // L- and S-steps from the old factors on the host while the device runs
      // the K range job
      wait_part();
      std::vector<HMat> XLX = galerkin_view(gx, r, r);
      HMat L = stream_L(aw, XLX);                          // (:1850-1854)
      HMat W1 = augment_w(L);
      HMat Sn = stream_S(S, XLX, WAW);                     // old Galerkin (:1859)
      fetch();
      range_finish(j, r, Xn + (size_t)r * C);             // Xn = [X | Xt]
      areset();
      const size_t xk = tn(Xn + (size_t)r * C, r, Kc, r, dev::OP_NONE);  // Xt^T K
      fetch();
      truncate_to(augmented_s(Sn, view(xk, r, r), L, W1), Xn, W1);
    }
    // X^T psi (old X) on the device during the host out-scattering
    const int r = rank;
    areset();
    const size_t xtp = tn_psi(X, r);
    fetch_part();
    out_scatter(Dout);
    // in-scattering (:1900-1940)
    const HMat MtDW = mtdw(Dvec, W);
    scatter_K(MtDW);
    copy_x_to_aug(r);
    dcopy(Kc, K, sizeof(double) * (size_t)C * r);
    RangeJob j = range_launch(K, r, X, r);
    // L- and S-steps on the host while the device runs the K range job
    wait_part();
    const HMat Xtpsi = view(xtp, r, nq);
    HMat L = titus::matmul(W, S, false, true);
    const HMat MX = titus::matmul(M, Xtpsi, false, true);
    for (int c = 0; c < r; ++c)
      for (int i = 0; i < nm; ++i) L(i, c) += dE * Dvec[i] * MX(i, c);
    HMat W1 = augment_w(L);
    const HMat add = titus::matmul(Xtpsi, MtDW);
    for (size_t i = 0; i < S.a.size(); ++i) S.a[i] += dE * add.a[i];
    fetch();
    range_finish(j, r, Xn + (size_t)r * C);
    areset();
    const size_t xk = tn(Xn + (size_t)r * C, r, Kc, r, dev::OP_NONE);
    fetch();
    const HMat Sn = augmented_s(S, view(xk, r, r), L, W1);
    dose_half(stopping[s]);                                // before the truncation (:1944-1950)
    truncate_to(Sn, Xn, W1);
    ranks.push_back(rank);                                 // rVec[2, n] (:1946)
  }

  void step(int s) {
    if (s != steps_done + 1) throw TitusError(TITUS_E_STATE, "steps must run in order");
    if (s >= ne) throw TitusError(TITUS_E_STATE, "step beyond the energy grid");
    timing.enabled = !group && timing.period > 0 && s % timing.period == 0;
    HostScope hs(timing, "host:step");
    if (adj) adjoint_checkpoint(*this, s);
    if (integrator == TITUS_INTEGRATOR_FIXED) step_fixed(s);
    else if (integrator == TITUS_INTEGRATOR_RK4) step_rk4(s);
    else if (integrator == TITUS_INTEGRATOR_PARALLEL) step_parallel(s);
    else step_bug(s);
    steps_done = s;
    for (double v : S.a)
      if (!std::isfinite(v)) throw TitusError(TITUS_E_NONFINITE, "non-finite coefficient matrix");
  }
};


// ---------------------------------------------------------------------------
namespace titus_host {

// Turns the per-operator host threads of a batch's contexts off for the
// batch's lifetime and restores each context's own setting afterwards (also
// when a solve fails).
struct HostParallelOff {
  std::vector<std::pair<titus_ctx*, bool>> saved;
  HostParallelOff(titus_ctx* const* ctxs, int n, bool on) {
    if (!on) return;
    for (int i = 0; i < n; ++i)
      if (ctxs[i]) {
        saved.emplace_back(ctxs[i], ctxs[i]->host_parallel);
        ctxs[i]->host_parallel = false;
      }
  }
  ~HostParallelOff() {
    for (auto& s : saved) s.first->host_parallel = s.second;
  }
};

template <class F>
inline int guarded(F&& f) {
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

inline void require(bool ok, const char* what) {
  if (!ok) throw TitusError(TITUS_E_ARG, what);
}

inline HMat copy_mat(const double* p, int m, int n) {
  HMat A(m, n);
  if (p) std::memcpy(A.a.data(), p, sizeof(double) * (size_t)m * n);
  return A;
}

}  // namespace titus_host
