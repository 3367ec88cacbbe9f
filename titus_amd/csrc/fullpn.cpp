// Full-rank P_N solver SolveFirstCollisionSource (code/SolverCSD.jl:465-586):
// context, block-sparse operator plan, energy step and its C-ABI
// (include/titus_hip.h).
//
// Per energy step, all on the context's stream:
//   bc       psi on the four edges = PsiBeam(energy[n], n-1)          (:528-537)
//   sweep    psi <- (psi - dE flux(psi)) / (1 + dE sigma_0), no 1/rho  (:549-553)
//            (the DLR path's sweep kernel with unit density and no restore)
//   collide  u  <- implicit scattering of the streamed u with the
//            in-scattering source of the new psi                     (:556-566)
//   dose     dose += dE (u[:,0] + uOUnc) SMid[n-1] / rho / w          (:575)
// The collided update dominates: its moment operators |Ax|, |Az|, Ax, Az are
// sparse (P_21: 25%, 3%, 1.4%, 0.7% nonzero).  The moments are reordered so
// that the operators become block-structured -- connected components of
// |Ax| + Ax + |Az| (the two parity classes of l + k), then components of |Az|
// (one per k within a class) -- and only the nonzero 4 x 16 blocks reach the
// matrix cores.  u lives in the reordered moment order on the device; the
// state is returned in the caller's order.
#include <functional>

#include "context.h"

namespace {

// u -> its moment order on the device: sort by (component of g1, component of
// g2, index), each component named by its smallest index.  Moment 0 stays
// first.
std::vector<int> moment_order(int nm, const std::vector<const double*>& g1,
                              const std::vector<const double*>& g2) {
  auto components = [nm](const std::vector<const double*>& ops) {
    std::vector<int> parent(nm);
    std::iota(parent.begin(), parent.end(), 0);
    std::function<int(int)> find = [&](int a) {
      while (parent[a] != a) a = parent[a] = parent[parent[a]];
      return a;
    };
    for (const double* A : ops)
      for (int j = 0; j < nm; ++j)
        for (int i = 0; i < nm; ++i)
          if (A[(size_t)j * nm + i] != 0.0) {
            int a = find(i), b = find(j);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
          }
    std::vector<int> rep(nm);
    for (int i = 0; i < nm; ++i) rep[i] = find(i);   // the smallest index of the component
    return rep;
  };
  const std::vector<int> k1 = components(g1), k2 = components(g2);
  std::vector<int> order(nm);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return std::make_tuple(k1[a], k2[a], a) < std::make_tuple(k1[b], k2[b], b);
  });
  return order;
}

// Block-sparse plan for `nbw` column blocks per wavefront (layout of
// dev::PnBlocks).  ops[o](k, m) in the device moment order.
struct HostPlan {
  int nbw = 0, ngroups = 0, ks = 1;
  std::vector<unsigned> mask;
  std::vector<int> start;
  std::vector<double> pk;
  long long nnz_blocks = 0;
};

HostPlan build_plan(int nm, int nbw, int ks, const std::vector<const double*>& ops,
                    const std::vector<int>& perm) {
  HostPlan h;
  const int kr = 4 * ks, nkb = (nm + kr - 1) / kr, nnb = (nm + 15) / 16;
  h.nbw = nbw;
  h.ks = ks;
  h.ngroups = (nnb + nbw - 1) / nbw;
  h.mask.assign((size_t)h.ngroups * nkb * 4, 0u);
  h.start.assign((size_t)h.ngroups * (nkb + 1), 0);
  auto val = [&](int o, int k, int m) {
    return (k < nm && m < nm) ? ops[o][(size_t)perm[m] * nm + perm[k]] : 0.0;
  };
  int blk = 0;
  for (int g = 0; g < h.ngroups; ++g) {
    for (int kb = 0; kb < nkb; ++kb) {
      h.start[(size_t)g * (nkb + 1) + kb] = blk;
      for (int o = 0; o < 4; ++o)
        for (int t = 0; t < nbw; ++t) {
          const int mb = g * nbw + t;
          if (mb >= nnb) continue;
          bool nz = false;
          for (int l = 0; l < 64 * ks && !nz; ++l)
            nz = val(o, kb * kr + l / 16, mb * 16 + l % 16) != 0.0;
          if (!nz) continue;
          h.mask[((size_t)g * nkb + kb) * 4 + o] |= 1u << t;
          for (int l = 0; l < 64 * ks; ++l) h.pk.push_back(val(o, kb * kr + l / 16, mb * 16 + l % 16));
          ++blk;
        }
    }
    h.start[(size_t)g * (nkb + 1) + nkb] = blk;
  }
  h.nnz_blocks = (long long)blk * ks;          // in 4 x 16 units
  // the kernel reads a K-block's full capacity (4 nbw blocks) from its start
  h.pk.resize(h.pk.size() + (size_t)4 * nbw * 64 * ks, 0.0);
  return h;
}

constexpr size_t kTuneMinWork = size_t(1) << 20;   // cells x moments worth timing plans for
std::mutex g_fcs_tune_mu;
struct FcsTuned { int nbw, ct, ks; dev::SweepPlan sweep; };
std::map<std::tuple<int, int, int, int>, FcsTuned> g_fcs_tuned;

}  // namespace

struct titus_fcs {
  int device = 0;
  hipStream_t st = nullptr;
  int nx = 0, ny = 0, C = 0, nm = 0, nq = 0, ne = 0, nsig = 0, nb = 0;
  double dx = 0, dy = 0;
  std::vector<double> dE, smid, bcs, sigma;
  std::vector<int> perm;            // device moment k is the caller's moment perm[k]
  int steps_done = 0;
  double *rho = nullptr, *irho = nullptr, *ones = nullptr, *u = nullptr, *un = nullptr;
  double *psi = nullptr, *psi2 = nullptr, *dose = nullptr, *u0p = nullptr, *omega = nullptr;
  double *bc_base = nullptr, *m0 = nullptr, *Mp = nullptr, *sig_d = nullptr, *one = nullptr;
  double* pk = nullptr;
  int *degp = nullptr, *bcells = nullptr, *start = nullptr;
  unsigned* mask = nullptr;
  unsigned char* restore = nullptr;
  dev::PnBlocks blocks;
  long long nnz_blocks = 0, dense_blocks = 0;
  bool tuned = false;
  dev::Grid g1{}, gr{};
  dev::SweepPlan plan;
  int G = 1;
  double collide_flops = 0.0;
  Timing timing;

  ~titus_fcs() {
    if (device >= 0) (void)hipSetDevice(device);
    if (st) (void)hipStreamSynchronize(st);
    void* ptrs[] = {rho, irho, ones, u, un, psi, psi2, dose, u0p, omega, bc_base, m0, Mp,
                    sig_d, one, pk, degp, bcells, start, mask, restore};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    timing.pending.clear();
    if (st) (void)hipStreamDestroy(st);
  }

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

  void upload_plan(const HostPlan& h) {
    if (pk) HIPCHK(hipFree(pk));
    if (mask) HIPCHK(hipFree(mask));
    if (start) HIPCHK(hipFree(start));
    pk = dalloc<double>(std::max<size_t>(h.pk.size(), 1));
    mask = dalloc<unsigned>(h.mask.size());
    start = dalloc<int>(h.start.size());
    if (!h.pk.empty())
      HIPCHK(hipMemcpy(pk, h.pk.data(), sizeof(double) * h.pk.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(mask, h.mask.data(), sizeof(unsigned) * h.mask.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(start, h.start.data(), sizeof(int) * h.start.size(), hipMemcpyHostToDevice));
    blocks.nm = nm;
    blocks.ks = h.ks;
    blocks.nkb = (nm + 4 * h.ks - 1) / (4 * h.ks);
    blocks.nnb = (nm + 15) / 16;
    blocks.nbw = h.nbw;
    blocks.ngroups = h.ngroups;
    blocks.mask = mask;
    blocks.start = start;
    blocks.pk = pk;
    nnz_blocks = h.nnz_blocks;
  }

  void collide(double dEs, const double* sig_row) {
    dev::launch_pn_collide(st, gr, blocks, u, un, psi2, nq, Mp, degp, sig_row, dEs);
  }

  void step(int s) {
    if (s != steps_done + 1 || s > ne - 1)
      throw TitusError(TITUS_E_STATE, "energy steps must be taken in order 1..ne-1");
    const double dEs = dE[s - 1];
    const double* sig = &sigma[(size_t)s * nsig];
    const double* sig_row = sig_d + (size_t)s * nsig;
    const double cq = 8.0 * (double)C * nq;
    timed("bc", 8.0 * nb * nq * 2, 0.0, [&] {
      dev::launch_bc(st, psi, dev::DT::F64, C, nq, bcells, nb, bc_base, bcs[s - 1], m0, u0p, G);
    });
    timed("sweep", 2.0 * cq, 6.0 * (double)C * nq, [&] {
      dev::launch_sweep(st, g1, psi, psi2, dev::DT::F64, nq, omega, dEs, 1.0 + dEs * sig[0],
                        restore, m0, u0p, plan);
    });
    // uOUnc is only set on interior cells (:565); the ghost cells' stay 0
    dev::launch_zero_rows(st, u0p, C, G, bcells, nb);
    timed("collide", 8.0 * (2.0 * C * nm) + cq, collide_flops, [&] { collide(dEs, sig_row); });
    const double w = (s == 1 || s == ne - 1) ? 2.0 : 1.0;
    timed("dose", 8.0 * (double)C * (3 + G), 3.0 * C * (1 + G), [&] {
      dev::launch_dose(st, gr, dose, un, C, 1, one, u0p, G, dEs * smid[s - 1] / w);
    });
    std::swap(u, un);
    std::swap(psi, psi2);
    steps_done = s;
  }

  // Plan choice: time every column-block width (and the sweep's launch
  // plans) on the context's own buffers for problems large enough to be
  // bandwidth or matrix-core bound; the heuristic plan otherwise.
  void tune(const std::vector<const double*>& ops) {
    const int nnb = (nm + 15) / 16;
    int widest = 1;
    while (widest < nnb && widest < dev::kPnMaxNbw) widest *= 2;
    plan = dev::sweep_default(g1, nq, dev::DT::F64);
    int best_nbw = std::min(widest, 8), best_ct = 1, best_ks = 1;
    tuned = false;
    const auto key = std::make_tuple(nx, ny, nm, nq);
    std::lock_guard<std::mutex> lock(g_fcs_tune_mu);
    if ((size_t)C * nm >= kTuneMinWork) {
      auto it = g_fcs_tuned.find(key);
      if (it != g_fcs_tuned.end()) {
        best_nbw = it->second.nbw;
        best_ct = it->second.ct;
        best_ks = it->second.ks;
        plan = it->second.sweep;
        tuned = true;
      } else {
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        auto time_it = [&](auto&& run) {
          run();
          HIPCHK(hipEventRecord(a, st));
          for (int rep = 0; rep < 3; ++rep) run();
          HIPCHK(hipEventRecord(b, st));
          HIPCHK(hipEventSynchronize(b));
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, a, b));
          return ms;
        };
        float best = 0.f;
        for (int nbw = 1; nbw <= widest; nbw *= 2)
          for (int ks = 1; ks <= (nbw <= 8 ? 2 : 1); ++ks) {
            upload_plan(build_plan(nm, nbw, ks, ops, perm));
            for (int ct = 1; ct <= (nbw <= 8 ? 2 : 1); ++ct) {
              blocks.ct = ct;
              const float ms = time_it([&] { collide(dE[0], sig_d + nsig); });
              if (!tuned || ms < best) {
                best = ms; best_nbw = nbw; best_ct = ct; best_ks = ks; tuned = true;
              }
            }
          }
        bool first = true;
        for (const dev::SweepPlan& c : dev::sweep_candidates(g1, nq, dev::DT::F64)) {
          const float ms = time_it([&] {
            dev::launch_sweep(st, g1, psi, psi2, dev::DT::F64, nq, omega, dE[0], 1.0, restore, m0,
                              u0p, c);
          });
          if (first || ms < best) { best = ms; plan = c; first = false; }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        g_fcs_tuned[key] = {best_nbw, best_ct, best_ks, plan};
      }
    }
    upload_plan(build_plan(nm, best_nbw, best_ks, ops, perm));
    blocks.ct = best_ct;
    G = dev::sweep_groups(g1, nq, plan);
    // the timing runs wrote into the zero buffers: clear them again
    HIPCHK(hipMemset(un, 0, sizeof(double) * (size_t)C * nm));
    HIPCHK(hipMemset(u, 0, sizeof(double) * (size_t)C * nm));
    HIPCHK(hipMemset(psi2, 0, sizeof(double) * (size_t)C * nq));
    HIPCHK(hipMemset(u0p, 0, sizeof(double) * (size_t)dev::kMaxSweepGroups * C));
  }
};

extern "C" {

int titus_fcs_create(titus_fcs** out, const titus_fcs_problem_v1* p, int32_t device) {
  return guarded([&] {
    require(out != nullptr && p != nullptr, "null argument");
    *out = nullptr;
    require(p->version == TITUS_ABI_VERSION, "problem struct version mismatch");
    require(p->nx >= 3 && p->ny >= 3, "need at least 3 cells per direction");
    require(p->nx == p->ny,
            "SolveFirstCollisionSource numbers cells (i-1)*nx + j (SolverCSD.jl:563): square grids only");
    require(p->dx > 0 && p->dy > 0, "cell widths must be positive");
    require(p->nm >= 1 && p->nq >= 1 && p->ne >= 3 && p->nsigma >= 1, "empty dimension");
    require(p->density && p->degree && p->Ax && p->Az && p->absAx && p->absAz && p->omega && p->M &&
                p->dE && p->stopping_mid && p->sigma && p->bcells && p->bc_base && p->bc_scale,
            "missing operand");
    require((long long)p->nx * p->ny < (1LL << 31), "grid too large");
    const int C = p->nx * p->ny;
    for (int c = 0; c < C; ++c)
      require(std::isfinite(p->density[c]) && p->density[c] > 0.0, "density must be positive and finite");
    for (int i = 0; i < p->nm; ++i)
      require(p->degree[i] >= 0 && p->degree[i] < p->nsigma, "moment degree outside sigma table");
    for (int b = 0; b < p->nb; ++b) require(p->bcells[b] >= 0 && p->bcells[b] < C, "ghost cell out of range");
    for (int s = 0; s < p->ne - 1; ++s)
      require(std::isfinite(p->dE[s]) && p->dE[s] > 0.0, "energy steps must be positive");
    if (p->psi0)
      for (size_t e = 0; e < (size_t)C * p->nq; ++e)
        require(std::isfinite(p->psi0[e]), "initial flux must be finite");
    // the ghost cells must be exactly the four edges: the kernels take the
    // interior from the grid, the reference from its loops (:528-537, :562-563)
    {
      std::vector<unsigned char> edge(C, 0);
      for (int b = 0; b < p->nb; ++b) edge[p->bcells[b]] = 1;
      for (int c = 0; c < C; ++c) {
        const int i = c / p->ny, j = c % p->ny;
        const bool on_edge = i == 0 || i == p->nx - 1 || j == 0 || j == p->ny - 1;
        require(edge[c] == (on_edge ? 1 : 0), "ghost cells must be the four edges of the grid");
      }
    }

    std::unique_ptr<titus_fcs> ctx(new titus_fcs());
    titus_fcs& x = *ctx;
    x.device = -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      throw TitusError(TITUS_E_NODEVICE, "no HIP device available");
    require(device >= 0 && device < ndev, "device index out of range");
    x.device = device;
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&x.st, hipStreamNonBlocking));
    x.nx = p->nx; x.ny = p->ny; x.C = C; x.nm = p->nm; x.nq = p->nq; x.ne = p->ne;
    x.nsig = p->nsigma; x.nb = p->nb; x.dx = p->dx; x.dy = p->dy;
    x.dE.assign(p->dE, p->dE + p->ne - 1);
    x.smid.assign(p->stopping_mid, p->stopping_mid + p->ne - 1);
    x.bcs.assign(p->bc_scale, p->bc_scale + p->ne - 1);
    x.sigma.assign(p->sigma, p->sigma + (size_t)p->ne * p->nsigma);
    const int nm = p->nm, nq = p->nq;

    // moment order and block plan
    const std::vector<const double*> ops = {p->absAx, p->absAz, p->Ax, p->Az};
    x.perm = moment_order(nm, {p->absAx, p->Ax, p->absAz}, {p->absAz});
    require(x.perm[0] == 0, "moment 0 must stay first");
    double nnz = 0.0;
    for (const double* A : ops)
      for (size_t e = 0; e < (size_t)nm * nm; ++e) nnz += A[e] != 0.0 ? 1.0 : 0.0;
    x.collide_flops = 2.0 * C * (nnz + (double)nm * nq);
    x.dense_blocks = 4LL * ((nm + 3) / 4) * ((nm + 15) / 16);

    const size_t cm = (size_t)C * nm, cq = (size_t)C * nq;
    x.rho = dalloc<double>(C);
    x.irho = dalloc<double>(C);
    x.ones = dalloc<double>(C);
    x.u = dalloc<double>(cm);
    x.un = dalloc<double>(cm);
    x.psi = dalloc<double>(cq);
    x.psi2 = dalloc<double>(cq);
    x.dose = dalloc<double>(C);
    x.u0p = dalloc<double>((size_t)dev::kMaxSweepGroups * C);
    x.omega = dalloc<double>((size_t)nq * 3);
    x.bc_base = dalloc<double>((size_t)std::max(1, p->nb) * nq);
    x.m0 = dalloc<double>(nq);
    x.Mp = dalloc<double>((size_t)nm * nq);
    x.sig_d = dalloc<double>((size_t)p->ne * p->nsigma);
    x.one = dalloc<double>(1);
    x.degp = dalloc<int>(nm);
    x.bcells = dalloc<int>(std::max(1, p->nb));
    x.restore = dalloc<unsigned char>(C);

    std::vector<double> ir(C), ones(C, 1.0);
    for (int c = 0; c < C; ++c) ir[c] = 1.0 / p->density[c];
    HIPCHK(hipMemcpy(x.rho, p->density, sizeof(double) * C, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.irho, ir.data(), sizeof(double) * C, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.ones, ones.data(), sizeof(double) * C, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.omega, p->omega, sizeof(double) * nq * 3, hipMemcpyHostToDevice));
    if (p->nb > 0)
      HIPCHK(hipMemcpy(x.bc_base, p->bc_base, sizeof(double) * (size_t)p->nb * nq, hipMemcpyHostToDevice));
    std::vector<double> m0(nq), Mp((size_t)nm * nq);
    for (int q = 0; q < nq; ++q) m0[q] = p->M[(size_t)q * nm + 0];
    for (int q = 0; q < nq; ++q)
      for (int k = 0; k < nm; ++k) Mp[(size_t)q * nm + k] = p->M[(size_t)q * nm + x.perm[k]];
    std::vector<int> degp(nm);
    for (int k = 0; k < nm; ++k) degp[k] = p->degree[x.perm[k]];
    const double one = 1.0;
    HIPCHK(hipMemcpy(x.m0, m0.data(), sizeof(double) * nq, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.Mp, Mp.data(), sizeof(double) * Mp.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.sig_d, p->sigma, sizeof(double) * x.sigma.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.one, &one, sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x.degp, degp.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
    if (p->nb > 0) HIPCHK(hipMemcpy(x.bcells, p->bcells, sizeof(int) * p->nb, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(x.restore, 0, C));
    HIPCHK(hipMemset(x.psi, 0, sizeof(double) * cq));
    HIPCHK(hipMemset(x.psi2, 0, sizeof(double) * cq));
    HIPCHK(hipMemset(x.u, 0, sizeof(double) * cm));
    HIPCHK(hipMemset(x.un, 0, sizeof(double) * cm));
    HIPCHK(hipMemset(x.dose, 0, sizeof(double) * C));
    HIPCHK(hipMemset(x.u0p, 0, sizeof(double) * dev::kMaxSweepGroups * C));
    // sweep on unit density (no 1/rho in the uncollided flux, :551); the
    // collided stencils and the dose use the density
    x.g1 = dev::Grid{p->nx, p->ny, C, p->dx, p->dy, x.ones, x.ones};
    x.gr = dev::Grid{p->nx, p->ny, C, p->dx, p->dy, x.rho, x.irho};
    x.tune(ops);
    if (p->psi0) HIPCHK(hipMemcpy(x.psi, p->psi0, sizeof(double) * cq, hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
    *out = ctx.release();
  });
}

int titus_fcs_step(titus_fcs* ctx, int32_t step) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    HIPCHK(hipSetDevice(ctx->device));
    ctx->step(step);
  });
}

int titus_fcs_solve(titus_fcs* ctx, int32_t last_step) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    HIPCHK(hipSetDevice(ctx->device));
    const int last = last_step <= 0 ? ctx->ne - 1 : std::min<int>(last_step, ctx->ne - 1);
    for (int s = ctx->steps_done + 1; s <= last; ++s) ctx->step(s);
    HIPCHK(hipStreamSynchronize(ctx->st));
  });
}

int titus_fcs_get_state(titus_fcs* ctx, double* u, double* dose, double* psi) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    titus_fcs& x = *ctx;
    HIPCHK(hipSetDevice(x.device));
    HIPCHK(hipStreamSynchronize(x.st));
    const size_t C = x.C;
    if (u) {
      std::vector<double> h(C * x.nm);
      HIPCHK(hipMemcpy(h.data(), x.u, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
      for (int k = 0; k < x.nm; ++k)
        std::memcpy(u + (size_t)x.perm[k] * C, h.data() + (size_t)k * C, sizeof(double) * C);
    }
    if (dose) HIPCHK(hipMemcpy(dose, x.dose, sizeof(double) * C, hipMemcpyDeviceToHost));
    if (psi) HIPCHK(hipMemcpy(psi, x.psi, sizeof(double) * C * x.nq, hipMemcpyDeviceToHost));
    std::vector<double> d(C);
    HIPCHK(hipMemcpy(d.data(), x.dose, sizeof(double) * C, hipMemcpyDeviceToHost));
    for (double v : d)
      if (!std::isfinite(v)) throw TitusError(TITUS_E_NONFINITE, "non-finite dose");
  });
}

int titus_fcs_set_timing(titus_fcs* ctx, int32_t on) {
  return guarded([&] {
    require(ctx != nullptr, "null context");
    HIPCHK(hipSetDevice(ctx->device));
    ctx->timing.resolve();
    ctx->timing.enabled = on != 0;
    ctx->timing.period = on != 0 ? 1 : 0;
    std::fill(ctx->timing.total_ms.begin(), ctx->timing.total_ms.end(), 0.0);
    std::fill(ctx->timing.total_bytes.begin(), ctx->timing.total_bytes.end(), 0.0);
    std::fill(ctx->timing.total_flops.begin(), ctx->timing.total_flops.end(), 0.0);
    std::fill(ctx->timing.launches.begin(), ctx->timing.launches.end(), 0);
  });
}

int titus_fcs_kernel_stats(const titus_fcs* ctx, const char* name, double* total_ms,
                           int64_t* launches, double* bytes, double* flops) {
  return guarded([&] {
    require(ctx != nullptr && name != nullptr, "null argument");
    titus_fcs& x = const_cast<titus_fcs&>(*ctx);
    HIPCHK(hipSetDevice(x.device));
    x.timing.resolve();
    auto it = x.timing.ids.find(name);
    const bool have = it != x.timing.ids.end();
    const int k = have ? it->second : 0;
    if (total_ms) *total_ms = have ? x.timing.total_ms[k] : 0.0;
    if (launches) *launches = have ? x.timing.launches[k] : 0;
    if (bytes) *bytes = have ? x.timing.total_bytes[k] : 0.0;
    if (flops) *flops = have ? x.timing.total_flops[k] : 0.0;
  });
}

int titus_fcs_get_plan(const titus_fcs* ctx, int32_t* values, int32_t n) {
  return guarded([&] {
    require(ctx != nullptr && (values != nullptr || n <= 0), "null argument");
    const int32_t v[7] = {ctx->blocks.nbw, ctx->blocks.ngroups, (int32_t)ctx->nnz_blocks,
                          (int32_t)ctx->dense_blocks, ctx->tuned ? 1 : 0, ctx->blocks.ct,
                          ctx->blocks.ks};
    for (int i = 0; i < n && i < 7; ++i) values[i] = v[i];
  });
}

void titus_fcs_destroy(titus_fcs* ctx) { delete ctx; }

}  // extern "C"
