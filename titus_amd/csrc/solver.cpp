// C-ABI of the DLRA energy loop (include/titus_hip.h): creation, stepping,
// batches, state and diagnostics.  The context and the step orchestration
// live in context.h; the adjoint in adjoint.cpp.
#include "context.h"

#include <condition_variable>
#include <functional>

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
  const bool lim = x.flux_order == 2;
  x.sweep_plan = dev::sweep_default(x.g, x.nq, x.tpsi);
  x.sweep_plan.limited = lim;
  x.sweep_tuned = false;
  if ((size_t)x.C * x.nq < kTuneMinWork) return;
  const auto key = std::make_tuple(x.nx, x.ny * x.nz + (x.nz > 1 ? 1 << 24 : 0), x.nq,
                                   (x.mixed ? 1 : 0) + (lim ? 2 : 0));
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
  for (dev::SweepPlan c : dev::sweep_candidates(x.g, x.nq, x.tpsi)) {
    if (lim && c.nontemporal) continue;
    c.limited = lim;
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
// Lockstep batch.  The contexts of a batch advance side by side, one host
// thread each, on the group's one stream.  A context's launches (kernels and
// copies) are recorded, not issued; when every live context of the group
// waits at a sync point (a fetch of device results, or leaving the group), the
// last to arrive issues the recorded launches of all of them: repeatedly the
// kernel that most contexts launch next, once for all those contexts
// (batch.h: k_many), so the launches of each context keep their order.  One
// copy of the launch items to the device precedes them; a stream sync ends
// the flush.  The host work of the contexts (ranks, completions, small
// algebra) runs in their threads between the syncs.
struct LockstepGroup {
  std::mutex m;
  std::condition_variable cv;
  int live = 0, arrived = 0;
  uint64_t gen = 0;
  bool failed = false;
  std::string error;
  std::vector<titus_ctx*> members;
  hipStream_t st = nullptr;
  unsigned char* h_items = nullptr;     // pinned staging of the launch items
  unsigned char* d_items = nullptr;
  size_t cap = 0;
  std::vector<unsigned char> staging;
  int64_t flushes = 0, launches = 0, ops = 0;
  int64_t longest = 0;                    // sum over flushes of the longest recorded list
  // Tile tuner of the group (a bucket of similar beamlets): for its first
  // kRounds rounds of kRoundFlushes flushes, round i gives every tuned
  // family its candidate i mod (count); every launch of a tuned family is
  // timed with events and its time divided by its work (16-cell tiles or
  // cell-ordinates); then each family keeps its fastest candidate.  The
  // candidates regroup the same tiles into workgroups (batch.h,
  // BatchTuning), so the tuner changes no result.
  static constexpr int kRoundFlushes = 4, kRounds = 10;
  dev::BatchTuning tune;
  bool tuning = true;
  int tuned_flushes = 0;
  const std::vector<int> cand_gemm{2, 4, 8}, cand_kstep{2, 4, 8};
  const std::vector<std::pair<int, int>> cand_sweep{{2, 0}, {4, 0}, {8, 0}, {2, 1}, {4, 1}};
  double t_sum[dev::TF_COUNT][8] = {}, w_sum[dev::TF_COUNT][8] = {};
  std::vector<hipEvent_t> evs;
  int ncand(int f) const {
    return f == 0 ? (int)cand_gemm.size() : f == 1 ? (int)cand_kstep.size() : (int)cand_sweep.size();
  }
  void set_round(int round) {
    tune.gemm_tpw = cand_gemm[round % cand_gemm.size()];
    tune.kstep_tpw = cand_kstep[round % cand_kstep.size()];
    tune.sweep_rows = cand_sweep[round % cand_sweep.size()].first;
    tune.sweep_nt = cand_sweep[round % cand_sweep.size()].second;
  }
  void finish_tuning() {
    int best[dev::TF_COUNT] = {0, 0, 0};
    for (int f = 0; f < dev::TF_COUNT; ++f) {
      double bt = -1.0;
      for (int c = 0; c < ncand(f); ++c)
        if (w_sum[f][c] > 0.0 && (bt < 0.0 || t_sum[f][c] / w_sum[f][c] < bt)) {
          bt = t_sum[f][c] / w_sum[f][c];
          best[f] = c;
        }
    }
    tune.gemm_tpw = cand_gemm[best[0]];
    tune.kstep_tpw = cand_kstep[best[1]];
    tune.sweep_rows = cand_sweep[best[2]].first;
    tune.sweep_nt = cand_sweep[best[2]].second;
    tuning = false;
  }
  double flush_s = 0.0, sync_s = 0.0;     // wall time in flushes, of it waiting for the stream
  std::vector<std::pair<int, size_t>> pending_timed;
  std::vector<double> tuned_work;

  LockstepGroup() { set_round(0); }
  ~LockstepGroup() {
    for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    if (h_items) (void)hipHostFree(h_items);
    if (d_items) (void)hipFree(d_items);
    if (st) (void)hipStreamDestroy(st);
  }
  void join(titus_ctx* c) {
    std::lock_guard<std::mutex> lk(m);
    ++live;
    if (std::find(members.begin(), members.end(), c) == members.end()) members.push_back(c);
  }
  void sync() {
    std::unique_lock<std::mutex> lk(m);
    if (failed) throw TitusError(TITUS_E_HIP, "lockstep batch: " + error);
    if (++arrived == live) {
      flush_locked();
    } else {
      const uint64_t g = gen;
      cv.wait(lk, [&] { return gen != g; });
    }
    if (failed) throw TitusError(TITUS_E_HIP, "lockstep batch: " + error);
  }
  // a context leaves (its solve ended or failed); its last launches stay
  // recorded until the next flush, which this call makes when everyone else
  // is already waiting (or nobody is left)
  void leave() {
    std::unique_lock<std::mutex> lk(m);
    --live;
    if (arrived == live) flush_locked();
  }
  void flush_locked() {
    try {
      flush_ops();
    } catch (const std::exception& e) {
      failed = true;
      error = e.what();
      (void)hipStreamSynchronize(st);
    }
    arrived = 0;
    ++gen;
    cv.notify_all();
  }
  struct Launch {
    dev::ManyFn fn; size_t off, isz; int count; unsigned blocks; size_t lds;
    int var = 0; const dev::ManyFn* alts = nullptr;
    int family = 0; double work = 0.0;
  };
  // every op of `r` (independent copies) in launches by kernel
  void plan_all(const dev::Recorder& r, std::vector<Launch>& plan) {
    std::vector<bool> done(r.ops.size(), false);
    for (size_t a = 0; a < r.ops.size(); ++a) {
      if (done[a]) continue;
      Launch L{r.ops[a].fn, (staging.size() + 15) & ~size_t(15), 0, 0, 0, 0, 0, r.ops[a].alts};
      staging.resize(L.off);
      for (size_t b = a; b < r.ops.size(); ++b) {
        const dev::RecOp& o = r.ops[b];
        if (done[b] || o.fn != L.fn) continue;
        done[b] = true;
        const size_t at = staging.size();
        staging.resize(at + o.size);
        std::memcpy(staging.data() + at, r.bytes.data() + o.off, o.size);
        ++L.count;
        L.isz = o.size;
        L.blocks = std::max(L.blocks, o.blocks);
        L.lds = std::max(L.lds, o.lds);
        L.var = std::max(L.var, o.var);
      }
      plan.push_back(L);
    }
  }
  void flush_ops() {
    const auto t0 = std::chrono::steady_clock::now();
    // the contexts' uploads since the last flush go to the device before the
    // first kernel, their result arenas to the host after the last one
    dev::Recorder pro, epi;
    for (titus_ctx* c : members) {
      if (c->upl_top > c->upl_sent) {
        dev::record_copy(pro, c->upl + c->upl_sent, c->h_upl_dev + c->upl_sent,
                         (c->upl_top - c->upl_sent) * sizeof(double));
        c->upl_sent = c->upl_top;
      }
      if (c->arena_top > c->fetched_top) {
        dev::record_copy(epi, c->h_arena_dev + c->fetched_top, c->arena + c->fetched_top,
                         (c->arena_top - c->fetched_top) * sizeof(double));
        c->fetched_top = c->arena_top;
      }
    }
    struct Head { dev::Recorder* r; size_t i; };
    std::vector<Head> heads;
    size_t lmax = 0;
    for (titus_ctx* c : members)
      if (!c->rec.ops.empty()) {
        heads.push_back({&c->rec, 0});
        lmax = std::max(lmax, c->rec.ops.size());
      }
    longest += (int64_t)lmax;
    std::vector<Launch> plan;
    std::vector<std::pair<dev::ManyFn, size_t>> next;
    staging.clear();
    plan_all(pro, plan);
    for (;;) {
      next.clear();
      // the next kernel: the one whose contexts have the most launches left
      // (the weighted majority merge of the recorded lists)
      for (const Head& h : heads) {
        if (h.i >= h.r->ops.size()) continue;
        const dev::ManyFn f = h.r->ops[h.i].fn;
        const size_t left = h.r->ops.size() - h.i;
        auto it = std::find_if(next.begin(), next.end(), [&](const auto& e) { return e.first == f; });
        if (it == next.end()) next.emplace_back(f, left);
        else it->second += left;
      }
      if (next.empty()) break;
      const auto best = *std::max_element(next.begin(), next.end(),
                                          [](const auto& a, const auto& b) { return a.second < b.second; });
      Launch L{best.first, (staging.size() + 15) & ~size_t(15), 0, 0, 0, 0, 0, nullptr};
      staging.resize(L.off);
      for (Head& h : heads) {
        if (h.i >= h.r->ops.size() || h.r->ops[h.i].fn != best.first) continue;
        const dev::RecOp& o = h.r->ops[h.i];
        const size_t at = staging.size();
        staging.resize(at + o.size);
        std::memcpy(staging.data() + at, h.r->bytes.data() + o.off, o.size);
        ++L.count;
        L.isz = o.size;
        L.blocks = std::max(L.blocks, o.blocks);
        L.lds = std::max(L.lds, o.lds);
        L.var = std::max(L.var, o.var);
        L.alts = o.alts;
        L.family = o.family;
        L.work += o.work;
        ++h.i;
        ++ops;
      }
      plan.push_back(L);
    }
    plan_all(epi, plan);
    if (!plan.empty()) {
      if (staging.size() > cap) {
        // the stream is idle here: every flush ends with a sync
        if (h_items) HIPCHK(hipHostFree(h_items));
        if (d_items) HIPCHK(hipFree(d_items));
        h_items = nullptr;
        d_items = nullptr;
        cap = std::max(staging.size(), 2 * cap);
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&h_items), cap, hipHostMallocDefault));
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&d_items), cap));
      }
      std::memcpy(h_items, staging.data(), staging.size());
      HIPCHK(hipMemcpyAsync(d_items, h_items, staging.size(), hipMemcpyHostToDevice, st));
      // events around the tuned launches while the tuner runs
      std::vector<std::pair<int, size_t>> timed;      // (family, first event)
      if (tuning) {
        size_t need = 0;
        for (const Launch& L : plan)
          if (L.family > 0) need += 2;
        while (evs.size() < need) {
          hipEvent_t e;
          HIPCHK(hipEventCreate(&e));
          evs.push_back(e);
        }
      }
      size_t ev_next = 0;
      for (const Launch& L : plan) {
        const bool time_it = tuning && L.family > 0;
        if (time_it) HIPCHK(hipEventRecord(evs[ev_next], st));
        // at most 65535 items per launch (the grid's y extent)
        for (int first = 0; first < L.count; first += 65535) {
          const int cnt = std::min(65535, L.count - first);
          const dev::ManyFn fn = L.alts ? L.alts[L.var] : L.fn;
          fn(st, d_items + L.off + (size_t)first * L.isz, cnt, L.blocks, L.lds);
          const hipError_t e = hipGetLastError();
          if (e != hipSuccess)
            throw TitusError(TITUS_E_HIP, std::string("batched launch: ") + hipGetErrorString(e));
          ++launches;
        }
        if (time_it) {
          HIPCHK(hipEventRecord(evs[ev_next + 1], st));
          timed.emplace_back(L.family, ev_next);
          tuned_work.push_back(L.work);
          ev_next += 2;
        }
      }
      pending_timed = std::move(timed);
    }
    const auto t1 = std::chrono::steady_clock::now();
    HIPCHK(hipStreamSynchronize(st));
    const auto t2 = std::chrono::steady_clock::now();
    if (tuning) {
      const int round = tuned_flushes / kRoundFlushes;
      for (size_t q = 0; q < pending_timed.size(); ++q) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, evs[pending_timed[q].second], evs[pending_timed[q].second + 1]));
        const int f = pending_timed[q].first - 1;
        const int c = round % ncand(f);
        t_sum[f][c] += ms;
        w_sum[f][c] += tuned_work[q];
      }
      pending_timed.clear();
      tuned_work.clear();
      ++tuned_flushes;
      if (tuned_flushes % kRoundFlushes == 0) {
        if (tuned_flushes / kRoundFlushes >= kRounds) finish_tuning();
        else set_round(tuned_flushes / kRoundFlushes);
      }
    }
    for (Head& h : heads) h.r->clear();
    ++flushes;
    flush_s += std::chrono::duration<double>(t2 - t0).count();
    sync_s += std::chrono::duration<double>(t2 - t1).count();
  }
};

void lockstep_sync(titus_ctx& x) { x.group->sync(); }


extern "C" {

int32_t titus_abi_version(void) { return TITUS_ABI_VERSION; }

const char* titus_last_error(void) { return g_last_error.c_str(); }

int titus_create(titus_ctx** out, const titus_problem_v1* p, int32_t device) {
  return guarded([&] {
    require(out != nullptr && p != nullptr, "null argument");
    *out = nullptr;
    require(p->version >= 1 && p->version <= TITUS_ABI_VERSION, "problem struct version mismatch");
    const int nz = p->version >= 2 ? p->nz : 1;
    require(p->nx >= 3 && p->ny >= 3 && (nz == 1 || nz >= 3), "need at least 3 cells per direction");
    if (nz > 1) require(p->dz > 0 && p->Ay && p->absAy, "3D grid needs dz, Ay and |Ay|");
    require(p->dx > 0 && p->dy > 0, "cell widths must be positive");
    require(p->nm >= 1 && p->nq >= 1 && p->ne >= 3 && p->nsigma >= 1, "empty dimension");
    require(p->nb >= 0 && p->nrestore >= 0 && p->ncand >= 0, "negative cell count");
    require(p->density && p->degree && p->Ax && p->Az && p->absAx && p->absAz && p->omega &&
                p->M && p->stopping && p->sigma && p->bcells && p->bc_base && p->bc_scale &&
                p->cand_cells, "missing operand");
    require(p->nrestore == 0 || p->restore_cells, "missing operand: restore_cells");
    require(std::isfinite(p->dE) && p->dE > 0.0, "energy step must be positive and finite");
    require(p->integrator == TITUS_INTEGRATOR_FIXED || p->integrator == TITUS_INTEGRATOR_BUG ||
                p->integrator == TITUS_INTEGRATOR_PARALLEL || p->integrator == TITUS_INTEGRATOR_RK4,
            "unknown integrator");
    require(p->precision == TITUS_PRECISION_FP64 || p->precision == TITUS_PRECISION_MIXED ||
                p->precision == TITUS_PRECISION_MIXED_FULL_FP32,
            "unknown precision form");
    require(p->r >= 1, "rank must be >= 1");
    const int flux_order = p->version >= 3 && p->flux_order != 0 ? p->flux_order : 1;
    require(flux_order == 1 || flux_order == 2, "flux order must be 1 (upwind) or 2 (limited)");
    require(flux_order == 1 || nz == 1, "the limited flux is 2D only");
    require((long long)p->nx * p->ny * nz < (1LL << 31), "grid too large");
    require(p->integrator != TITUS_INTEGRATOR_RK4 || nz == 1, "the RK4 variant is 2D only");
    const int C = p->nx * p->ny * nz;
    for (int c = 0; c < C; ++c)
      require(std::isfinite(p->density[c]) && p->density[c] > 0.0, "density must be positive and finite");
    for (int i = 0; i < p->nm; ++i)
      require(p->degree[i] >= 0 && p->degree[i] < p->nsigma, "moment degree outside sigma table");
    for (int b = 0; b < p->nb; ++b) require(p->bcells[b] >= 0 && p->bcells[b] < C, "ghost cell out of range");
    for (int b = 0; b < p->nrestore; ++b)
      require(p->restore_cells[b] >= 0 && p->restore_cells[b] < C, "restore cell out of range");
    for (int b = 0; b < p->ncand; ++b) require(p->cand_cells[b] >= 0 && p->cand_cells[b] < C, "candidate out of range");
    if (p->version >= 3 && p->psi0)
      for (size_t e = 0; e < (size_t)C * p->nq; ++e)
        require(std::isfinite(p->psi0[e]), "initial flux must be finite");

    std::unique_ptr<titus_ctx> ctx(new titus_ctx());
    titus_ctx& x = *ctx;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw TitusError(TITUS_E_NODEVICE, "no HIP device available");
    require(device >= 0 && device < ndev, "device index out of range");
    x.device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&x.st, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&x.ev_part, hipEventDisableTiming));
    x.nx = p->nx; x.ny = p->ny; x.C = C; x.nm = p->nm; x.nq = p->nq; x.ne = p->ne;
    x.nsig = p->nsigma; x.integrator = p->integrator; x.precision = p->precision;
    x.mixed = p->precision != TITUS_PRECISION_FP64;
    x.full32 = p->precision == TITUS_PRECISION_MIXED_FULL_FP32;
    x.tpsi = x.mixed ? dev::DT::F32 : dev::DT::F64;
    x.dx = p->dx; x.dy = p->dy; x.dE = p->dE; x.theta = p->theta; x.theta_index = p->theta_index;
    x.rank_tol = p->rank_tol > 0.0 ? p->rank_tol : 1e-7;
    x.rank = p->r;
    const bool adaptive = p->integrator == TITUS_INTEGRATOR_BUG || p->integrator == TITUS_INTEGRATOR_PARALLEL;
    x.rank_cap = adaptive ? p->rank_cap : p->r;
    require(x.rank_cap >= x.rank, "rank cap below the initial rank");
    x.capx = adaptive ? 2 * x.rank_cap : x.rank;
    require(x.capx <= p->nm || !adaptive, "2 * rank cap exceeds the moment count");
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
    x.flux_order = flux_order;
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
    {
      double hmin = std::min(p->dx, p->dy), irmax = 0.0;
      if (nz > 1) hmin = std::min(hmin, p->dz);
      for (int c = 0; c < C; ++c) irmax = std::max(irmax, 1.0 / p->density[c]);
      x.cfl = p->dE / hmin * irmax;
    }

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
    if (p->integrator == TITUS_INTEGRATOR_PARALLEL || p->integrator == TITUS_INTEGRATOR_RK4)
      x.Kc = dalloc<double>(cx);
    x.Q1 = dalloc<double>(cx);
    x.Q2 = dalloc<double>(cx);
    const int wmax = std::max(x.capx, p->nq);
    x.partial = dalloc<double>(dev::tn_partial_size());
    const size_t w2 = (size_t)wmax * wmax;
    const size_t ncand_used = std::min<size_t>(p->ncand, (size_t)2 * x.capx + 64);
    // largest phase: Galerkin of the augmented basis plus one range job
    x.arena_n = 8 * w2 + 10 * (size_t)x.capx * x.capx + ncand_used * 2 * x.capx + 256 +
                (size_t)x.nops() * x.capx * x.capx + (size_t)p->nm * x.capx;
    x.arena = dalloc<double>(x.arena_n);
    x.h_arena = halloc<double>(x.arena_n);
    x.upl_n = 64 * w2 + 8192;
    x.upl = dalloc<double>(x.upl_n);
    x.h_upl = halloc<double>(x.upl_n);
    x.setup_angular();
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
    // the RK4 variant has no sweep: its analytic flux writes one moment partial
    if (p->integrator != TITUS_INTEGRATOR_RK4) tune_sweep(x);
    x.u0_groups = p->integrator == TITUS_INTEGRATOR_RK4 ? 1 : dev::sweep_groups(x.g, p->nq, x.sweep_plan);
    HIPCHK(hipMemset(x.u0p, 0, sizeof(double) * dev::kMaxSweepGroups * C));
    HIPCHK(hipMemset(x.psi2, 0, psz * (size_t)C * p->nq));
    HIPCHK(hipMemset(x.X, 0, sizeof(double) * cx));
    if (p->version >= 3 && p->psi0) {
      // initial flux (SetupIC) in the storage type, and its l = 0 moment as
      // partial 0 of the moment sums the first dose half reads
      std::vector<double> u0(C, 0.0);
      if (x.mixed) {
        std::vector<float> f((size_t)C * p->nq);
        for (size_t e = 0; e < f.size(); ++e) f[e] = (float)p->psi0[e];
        for (int q = 0; q < p->nq; ++q)
          for (int c = 0; c < C; ++c) u0[c] += (double)f[(size_t)q * C + c] * m0[q];
        HIPCHK(hipMemcpy(x.psi, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
      } else {
        for (int q = 0; q < p->nq; ++q)
          for (int c = 0; c < C; ++c) u0[c] += p->psi0[(size_t)q * C + c] * m0[q];
        HIPCHK(hipMemcpy(x.psi, p->psi0, sizeof(double) * (size_t)C * p->nq, hipMemcpyHostToDevice));
      }
      HIPCHK(hipMemcpy(x.u0p, u0.data(), sizeof(double) * C, hipMemcpyHostToDevice));
    }

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
  if (batch_has_duplicates(ctxs, n)) {
    g_last_error = "a context appears twice in the batch";
    return TITUS_E_ARG;
  }
  int nt = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = std::max(1, std::min<int>(nt, n));
  std::vector<int> codes(n, TITUS_OK);
  std::vector<std::string> msgs(n);
  std::atomic<int> next{0};
  // the contexts run side by side: no per-operator host threads inside them
  // during the batch; each context's own setting returns afterwards
  HostParallelOff hp(ctxs, n, nt > 1);
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

} // extern "C"

int lockstep_run(titus_ctx** ctxs, int n, int width, int groups, int32_t* status, int64_t* stats,
                 int32_t* tiles, const std::function<void(titus_ctx&, int)>& work) {
  if (n == 0) return TITUS_OK;
  const int w = std::max(1, std::min<int>(width > 0 ? width : n, n));
  std::vector<int> codes(n, TITUS_OK);
  std::vector<std::string> msgs(n);
  // `groups` groups of about w / groups contexts, each on its own stream: the
  // host work of one group overlaps the device work of the others
  const int G = std::max(1, std::min(groups > 0 ? groups : 1, w));
  std::vector<std::unique_ptr<LockstepGroup>> grps;
  for (int k = 0; k < G; ++k) grps.emplace_back(new LockstepGroup());
  std::vector<int> gof(n, 0);
  const int dev_id = ctxs[0]->device;
  int rc = guarded([&] {
    HIPCHK(hipSetDevice(dev_id));
    for (auto& gp : grps) HIPCHK(hipStreamCreateWithFlags(&gp->st, hipStreamNonBlocking));
    for (int i = 0; i < n; ++i) {
      titus_ctx* c = ctxs[i];
      // the context's own queued work (creation) is done before the group's
      HIPCHK(hipStreamSynchronize(c->st));
      HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->h_arena_dev), c->h_arena, 0));
      HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->h_upl_dev), c->h_upl, 0));
    }
  });
  if (rc != TITUS_OK) return rc;
  HostParallelOff hp(ctxs, n, true);
  // the first w contexts join before any thread starts, so that they start
  // their first steps together and reach every sync point together (a
  // context that joined after the others had flushed would stay some phases
  // apart, and its launches would not share the others')
  // groups are buckets of similar shape: the first w contexts sorted by
  // moment count and rank capacity, dealt out in contiguous runs, so that a
  // group's launches share kernel variants and tile counts
  std::vector<int> start(w);
  std::iota(start.begin(), start.end(), 0);
  // (an adjoint batch first sorts by the length of the last checkpoint
  // segment, which it recomputes and reverses first: contexts with equal
  // lengths run their recomputations and reverse sweeps in phase)
  auto key = [&](int i) {
    const titus_ctx* c = ctxs[i];
    return std::make_tuple(adjoint_phase_key(*c), c->nm, c->capx);
  };
  std::stable_sort(start.begin(), start.end(), [&](int a, int b) { return key(a) < key(b); });
  std::vector<int> tgroup(w);
  for (int t = 0; t < w; ++t) {
    tgroup[t] = (int)((int64_t)t * G / w);
    gof[start[t]] = tgroup[t];
    grps[tgroup[t]]->join(ctxs[start[t]]);
  }
  std::atomic<int> next{w};
  auto worker = [&](int t) {
    (void)hipSetDevice(dev_id);
    LockstepGroup& grp = *grps[tgroup[t]];
    for (int i = start[t];;) {
      if (i >= n) return;
      titus_ctx* c = ctxs[i];
      if (i >= w) {
        gof[i] = tgroup[t];
        grp.join(c);
      }
      c->group = &grp;
      dev::t_recorder = &c->rec;
      dev::t_tune = &grp.tune;
      codes[i] = guarded([&] { work(*c, i); });
      if (codes[i] != TITUS_OK) msgs[i] = g_last_error;
      dev::t_recorder = nullptr;
      dev::t_tune = nullptr;
      c->group = nullptr;
      grp.leave();
      i = next.fetch_add(1);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < w; ++t) pool.emplace_back(worker, t);
  worker(0);
  for (auto& th : pool) th.join();
  // every context has left; the last leave of each group flushed what remained
  if (stats) {
    std::fill(stats, stats + 6, int64_t(0));
    for (auto& gp : grps) {
      stats[0] += gp->flushes;
      stats[1] += gp->launches;
      stats[2] += gp->ops;
      stats[3] += (int64_t)(gp->flush_s * 1e9);
      stats[4] += (int64_t)(gp->sync_s * 1e9);
      stats[5] += gp->longest;
    }
  }
  if (tiles)
    for (int k = 0; k < G; ++k) {
      const dev::BatchTuning& t = grps[k]->tune;
      tiles[4 * k] = t.gemm_tpw;
      tiles[4 * k + 1] = t.kstep_tpw;
      tiles[4 * k + 2] = t.sweep_rows;
      tiles[4 * k + 3] = t.sweep_nt;
    }
  int first = TITUS_OK;
  for (int i = 0; i < n; ++i) {
    const LockstepGroup& grp = *grps[gof[i]];
    if (codes[i] == TITUS_OK && grp.failed) {
      codes[i] = TITUS_E_HIP;
      msgs[i] = "lockstep batch: " + grp.error;
    }
    if (status) status[i] = codes[i];
    if (first == TITUS_OK && codes[i] != TITUS_OK) {
      first = codes[i];
      g_last_error = "solve " + std::to_string(i) + ": " + msgs[i];
    }
  }
  return first;
}


extern "C" {

int titus_solve_batch_lockstep(titus_ctx** ctxs, int32_t n, int32_t width, int32_t groups,
                               int32_t* status, int64_t* stats, int32_t* tiles) {
  if (!ctxs || n < 0) {
    g_last_error = "bad argument";
    return TITUS_E_ARG;
  }
  if (batch_has_duplicates(ctxs, n)) {
    g_last_error = "a context appears twice in the batch";
    return TITUS_E_ARG;
  }
  for (int i = 0; i < n; ++i) {
    titus_ctx* c = ctxs[i];
    std::string why;
    if (!c) why = "null context";
    else if (c->device != ctxs[0]->device) why = "contexts on different devices";
    else if (c->integrator != TITUS_INTEGRATOR_FIXED && c->integrator != TITUS_INTEGRATOR_BUG &&
             c->integrator != TITUS_INTEGRATOR_PARALLEL)
      why = "integrator without a batched form (RK4)";
    else if (c->nd != 2) why = "3D grid (no batched sweep)";
    else if (c->flux_order != 1) why = "limited flux (no batched sweep)";
    else if (c->tape) why = "adjoint recomputation in progress";
    if (!why.empty()) {
      g_last_error = "lockstep batch, context " + std::to_string(i) + ": " + why;
      return TITUS_E_ARG;
    }
  }
  return lockstep_run(ctxs, n, width, groups, status, stats, tiles, [](titus_ctx& c, int) {
    const int last = c.integrator == TITUS_INTEGRATOR_BUG ? c.ne - 2 : c.ne - 1;
    for (int s = c.steps_done + 1; s <= last; ++s) c.step(s);
  });
}

int titus_set_analytic_beam(titus_ctx* ctx, const titus_beam_v1* b) {
  return guarded([&] {
    require(ctx != nullptr && b != nullptr, "null argument");
    require(b->xmid && b->ymid && b->tau, "missing operand");
    if (ctx->integrator != TITUS_INTEGRATOR_RK4)
      throw TitusError(TITUS_E_ARG, "analytic beam: RK4 integrator only");
    if (ctx->steps_done != 0) throw TitusError(TITUS_E_STATE, "the beam must be set before the first step");
    for (double v : {b->x0, b->y0, b->Omega1, b->Omega3})
      require(std::isfinite(v), "beam parameters must be finite");
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->xmid_d) {
      ctx->xmid_d = dalloc<double>(ctx->nx);
      ctx->ymid_d = dalloc<double>(ctx->ny);
    }
    HIPCHK(hipMemcpy(ctx->xmid_d, b->xmid, sizeof(double) * ctx->nx, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->ymid_d, b->ymid, sizeof(double) * ctx->ny, hipMemcpyHostToDevice));
    ctx->tau.assign(b->tau, b->tau + ctx->ne);
    ctx->beam_x0 = b->x0; ctx->beam_y0 = b->y0;
    ctx->beam_om1 = b->Omega1; ctx->beam_om3 = b->Omega3;
    // int sigma starts at dE sigma_0(E_1) (SolverCSD.jl:1532)
    ctx->int_sigma = ctx->dE * ctx->sigma[0];
    ctx->beam_set = true;
  });
}

int titus_set_rejection(titus_ctx* ctx, double c_eta) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    require(std::isfinite(c_eta) && c_eta >= 0.0, "c_eta must be finite and >= 0");
    if (ctx->integrator != TITUS_INTEGRATOR_BUG)
      throw TitusError(TITUS_E_ARG, "rejection: rank-adaptive BUG integrator only");
    if (ctx->steps_done != 0) throw TitusError(TITUS_E_STATE, "rejection must be set before the first step");
    HIPCHK(hipSetDevice(ctx->device));
    if (c_eta > 0.0 && !ctx->X_start) {
      const size_t psz = ctx->mixed ? sizeof(float) : sizeof(double);
      ctx->X_start = dalloc<double>((size_t)ctx->C * ctx->capx);
      ctx->psi_start = dalloc<unsigned char>((size_t)ctx->C * ctx->nq * psz);
      ctx->dose_start = dalloc<double>(ctx->C);
      ctx->u0_start = dalloc<double>((size_t)dev::kMaxSweepGroups * ctx->C);
    }
    ctx->c_eta = c_eta;
  });
}

int titus_get_rejections(const titus_ctx* ctx, int64_t* count) {
  return guarded([&] {
    require(ctx != nullptr && count != nullptr, "null argument");
    *count = ctx->rejections;
  });
}

int titus_get_cfl(const titus_ctx* ctx, double* cfl) {
  return guarded([&] {
    require(ctx != nullptr && cfl != nullptr, "null argument");
    *cfl = ctx->cfl;
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

int titus_get_decisions(const titus_ctx* ctx, uint64_t* hash, int64_t* count) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    if (hash) *hash = ctx->decision_hash;
    if (count) *count = ctx->decisions;
  });
}

void titus_destroy(titus_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->st) (void)hipStreamSynchronize(ctx->st);
  delete ctx;
}

}  // extern "C"
